"""Builds the Fortran check program of the iso_c_binding layer:
include/hipnegf_cudautils.f90 + tests/fortran/blockops_check.f90, linked
against libnegf_amd/_lib/libhipnegf.so, into tests/fortran/_build/.

    python tests/fortran/build.py [outdir]

Returns without building when amdflang is absent."""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
FC = os.environ.get("FC_AMD", shutil.which("amdflang") or "/opt/rocm/llvm/bin/amdflang")


def build(outdir: str | None = None) -> str | None:
    if not os.path.exists(FC):
        return None
    out = outdir or os.path.join(HERE, "_build")
    os.makedirs(out, exist_ok=True)
    libdir = os.path.join(ROOT, "libnegf_amd", "_lib")
    mod_o = os.path.join(out, "hipnegf_cudautils.o")
    exe = os.path.join(out, "blockops_check")
    run = lambda cmd: subprocess.run(cmd, check=True, cwd=out, stdout=subprocess.PIPE,
                                     stderr=subprocess.STDOUT, text=True)
    run([FC, "-O1", "-c", os.path.join(ROOT, "include", "hipnegf_cudautils.f90"), "-o", mod_o])
    rpath = os.path.relpath(libdir, out) if outdir is None else libdir
    rpath = "$ORIGIN/" + rpath if outdir is None else rpath
    run([FC, "-O1", "-Wno-interoperability", os.path.join(HERE, "blockops_check.f90"), mod_o,
         "-L" + libdir, "-lhipnegf", "-Wl,-rpath," + rpath, "-o", exe])
    return exe


def build_mock(outdir: str) -> str | None:
    """The same program linked against the host stand-in mock_blockops.c
    (CPU tests of the interfaces only)."""
    if not os.path.exists(FC):
        return None
    os.makedirs(outdir, exist_ok=True)
    run = lambda cmd: subprocess.run(cmd, check=True, cwd=outdir, stdout=subprocess.PIPE,
                                     stderr=subprocess.STDOUT, text=True)
    lib = os.path.join(outdir, "libhipnegf_mock.so")
    run(["gcc", "-O1", "-fPIC", "-shared", os.path.join(HERE, "mock_blockops.c"), "-o", lib, "-lm"])
    mod_o = os.path.join(outdir, "hipnegf_cudautils.o")
    run([FC, "-O1", "-c", os.path.join(ROOT, "include", "hipnegf_cudautils.f90"), "-o", mod_o])
    exe = os.path.join(outdir, "blockops_check_mock")
    run([FC, "-O1", "-Wno-interoperability", os.path.join(HERE, "blockops_check.f90"), mod_o,
         "-L" + outdir, "-lhipnegf_mock", "-Wl,-rpath," + outdir, "-o", exe])
    return exe


if __name__ == "__main__":
    print(build(sys.argv[1] if len(sys.argv) > 1 else None))
