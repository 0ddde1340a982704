"""Python binding of the serving loop (hipnegf_server_* in include/hipnegf.h).

A ``Server`` owns W worker threads inside libhipnegf.so, each with its own
engine and HIP stream.  ``submit_kslab`` queues a transmission request on a
slab generated in HBM from (seed, n, nbl, k); ``submit`` queues one on a
resident ``DeviceStructure``.  ``wait`` blocks for a request's result.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np

from . import hip


class ServerOpts(C.Structure):
    _fields_ = [("workers", C.c_int), ("policy", C.c_int), ("max_chunk", C.c_int),
                ("reserved", C.c_int), ("chunk_flops", C.c_double), ("mem_budget", C.c_longlong),
                ("worker_workspace", C.c_longlong), ("sgf_cache", C.c_longlong),
                ("keep_structures", C.c_int), ("reserved2", C.c_int)]


FCFS, LAS = 0, 1


def _lib():
    L = hip.load()
    if not getattr(L, "_serve_bound", False):
        vp, i, d, u64 = C.c_void_p, C.c_int, C.c_double, C.c_ulonglong
        L.hipnegf_server_create.argtypes = [i, C.POINTER(ServerOpts), C.POINTER(vp)]
        L.hipnegf_server_destroy.argtypes = [vp]
        L.hipnegf_server_submit_kslab.argtypes = [vp, u64, i, i, d, d, i, i, vp, C.POINTER(hip.Opts),
                                                  C.POINTER(u64)]
        L.hipnegf_server_submit.argtypes = [vp, vp, i, i, vp, C.POINTER(hip.Opts), C.POINTER(u64)]
        L.hipnegf_server_poll.argtypes = [vp, u64, C.POINTER(i)]
        L.hipnegf_server_wait.argtypes = [vp, u64, vp, vp, vp, vp]
        L.hipnegf_server_stats.argtypes = [vp, vp]
        L.hipnegf_server_cache_stats.argtypes = [vp, vp]
        L.hipnegf_struct_generate_kslab.argtypes = [vp, vp, i, u64, d, d, i]
        L.hipnegf_struct_get_block.argtypes = [vp, i, i, i, vp]
        L._serve_bound = True
    return L


class Server:
    def __init__(self, device: int = 0, workers: int = 2, policy: int = FCFS, max_chunk: int = 0,
                 chunk_flops: float = 0.0, mem_budget: int = 0, worker_workspace: int = 0,
                 sgf_cache: int = 0, keep_structures: bool = True):
        L = _lib()
        o = ServerOpts(workers, policy, max_chunk, 0, chunk_flops, mem_budget, worker_workspace,
                       sgf_cache, 1 if keep_structures else 0, 0)
        h = C.c_void_p()
        hip.check(L.hipnegf_server_create(device, C.byref(o), C.byref(h)))
        self._h = h
        self._meta = {}

    def close(self):
        if getattr(self, "_h", None):
            hip.check(_lib().hipnegf_server_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit_kslab(self, seed: int, n: int, nbl: int, E: np.ndarray, k=(0.0, 0.0),
                     overlap: bool = True, delta: float = 1e-5, precision: int = hip.FP64,
                     ns_steps: int = 0) -> int:
        """precision=hip.MIXED: FP32 sweep with ns_steps FP64 Newton-Schulz
        steps per block inverse (the decimation stays FP64)."""
        E = np.ascontiguousarray(E, dtype=np.float64)
        o = hip._opts(delta, 1, 2, precision, 0, ns_steps=ns_steps)
        rid = C.c_ulonglong()
        hip.check(_lib().hipnegf_server_submit_kslab(self._h, seed, n, nbl, float(k[0]), float(k[1]),
                                                     1 if overlap else 0, len(E), hip._ptr(E),
                                                     C.byref(o), C.byref(rid)))
        self._meta[rid.value] = (len(E), 2)
        return rid.value

    def submit(self, ds: "hip.DeviceStructure", E: np.ndarray, slot: int = 0, delta: float = 1e-5,
               ni: int = 1, nf: int = 2) -> int:
        E = np.ascontiguousarray(E, dtype=np.float64)
        o = hip._opts(delta, ni, nf, hip.FP64, 0)
        rid = C.c_ulonglong()
        hip.check(_lib().hipnegf_server_submit(self._h, ds.handle, slot, len(E), hip._ptr(E),
                                               C.byref(o), C.byref(rid)))
        self._meta[rid.value] = (len(E), max(ds.template.ncont, 1))
        return rid.value

    def poll(self, rid: int) -> bool:
        d = C.c_int()
        hip.check(_lib().hipnegf_server_poll(self._h, rid, C.byref(d)))
        return bool(d.value)

    def wait(self, rid: int) -> dict:
        nE, nc = self._meta.pop(rid)
        T = np.zeros(nE)
        ncyc = np.zeros((nE, nc), dtype=np.int32)
        st = np.zeros(nE, dtype=np.int32)
        t = np.zeros(4)
        hip.check(_lib().hipnegf_server_wait(self._h, rid, hip._ptr(T), hip._ptr(ncyc), hip._ptr(st),
                                             hip._ptr(t)))
        return dict(T=T, ncyc=ncyc, status=st, t_queue=t[0], latency=t[1], t_gpu=t[2],
                    chunks=int(t[3]))

    def cache_stats(self) -> dict:
        out = np.zeros(4, dtype=np.int64)
        hip.check(_lib().hipnegf_server_cache_stats(self._h, hip._ptr(out)))
        return dict(structure_hits=int(out[0]), resident_structures=int(out[1]),
                    sgf_hits=int(out[2]), sgf_stores=int(out[3]))

    def stats(self) -> dict:
        out = np.zeros(4, dtype=np.int64)
        hip.check(_lib().hipnegf_server_stats(self._h, hip._ptr(out)))
        return dict(submitted=int(out[0]), completed=int(out[1]), running=int(out[2]),
                    queued=int(out[3]))


def kslab_structure(eng: "hip.Engine", seed: int, n: int, nbl: int, k=(0.0, 0.0),
                    overlap: bool = True, download: bool = False):
    """Generate the slab of (seed, n, nbl, k) on the device.  Returns the
    DeviceStructure (one slot) and, with download=True, the same blocks as a
    host BlockStructure (for checks against the oracle)."""
    from .structure import BlockStructure, Contact
    L = _lib()
    smode = (hip.S_DIAG | hip.S_OFFDIAG | hip.S_CONTACT) if overlap else 0
    ds = hip.DeviceStructure.empty(n=n, nbl=nbl, cblk=(nbl, 1), smode=smode)
    hip.check(L.hipnegf_struct_generate_kslab(eng.handle, ds.handle, 0, seed, float(k[0]),
                                              float(k[1]), 1 if overlap else 0))
    if not download:
        return ds, None

    def get(kind, idx):
        a = np.zeros(n * n, dtype=np.complex128)
        hip.check(L.hipnegf_struct_get_block(ds.handle, 0, kind, idx, hip._ptr(a)))
        return a.reshape((n, n), order="F")

    Hd = [get(hip.HD, i) for i in range(nbl)]
    Hu = [get(hip.HU, i) for i in range(nbl - 1)]
    Hl = [get(hip.HL, i) for i in range(nbl - 1)]
    Sd = [get(hip.SD, i) for i in range(nbl)] if overlap else None
    Su = [get(hip.SU, i) for i in range(nbl - 1)] if overlap else None
    Sl = [get(hip.SL, i) for i in range(nbl - 1)] if overlap else None
    cs = []
    for c, cb in enumerate((nbl, 1)):
        kw = dict(cblk=cb, H00=get(hip.H00, c), H01=get(hip.H01, c), HMC=get(hip.HMC, c))
        if overlap:
            kw.update(S00=get(hip.S00, c), S01=get(hip.S01, c), SMC=get(hip.SMC, c))
        cs.append(Contact(**kw))
    return ds, BlockStructure(Hd=Hd, Hu=Hu, Hl=Hl, contacts=cs, Sd=Sd, Su=Su, Sl=Sl,
                              name=f"generated kslab seed={seed} n={n} nbl={nbl}")
