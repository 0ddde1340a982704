"""ctypes binding of libhipnegf.so (the C ABI declared in include/hipnegf.h).

Loading is explicit and fails loudly: there is no CPU fallback anywhere in
this package.  The library must have been built in-tree
(``python -m libnegf_amd.build``).
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional, Sequence

import numpy as np

from .structure import BlockStructure

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libhipnegf.so")
_LIB = None

# block kinds / modes (hipnegf.h)
HD, SD, HU, SU, HL, SL, H00, S00, H01, S01, HMC, SMC = range(12)
HSS, SSS, HS1, SS1, H1S, S1S = range(12, 18)  # contact surface blocks
S_DIAG, S_OFFDIAG, S_CONTACT = 1, 2, 4
FP64, MIXED = 0, 1
ST_OK, ST_SINGULAR, ST_NOT_CONVERGED, ST_NONFINITE = 0, 1, 2, 4


class HipNegfError(RuntimeError):
    pass


class Opts(C.Structure):
    _fields_ = [("delta", C.c_double), ("ni", C.c_int), ("nf", C.c_int),
                ("precision", C.c_int), ("max_batch", C.c_int), ("ns_steps", C.c_int),
                ("polish", C.c_int), ("eps_drop", C.c_double), ("flags", C.c_int),
                ("lanes", C.c_int)]
NO_SHARED_DECIMATION = 1
GRAPH = 2
SGF_LOAD = 4
SGF_STORE = 8
NO_FUSED = 16


def load():
    """Load libhipnegf.so; raises if it has not been built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise HipNegfError(f"{LIB_PATH} is missing: build it with `python -m libnegf_amd.build` "
                           "(the HIP engine has no CPU fallback)")
    L = C.CDLL(LIB_PATH)
    vp, i, d, ll = C.c_void_p, C.c_int, C.c_double, C.c_longlong
    L.hipnegf_last_error.restype = C.c_char_p
    L.hipnegf_engine_create.argtypes = [i, C.POINTER(vp)]
    L.hipnegf_engine_destroy.argtypes = [vp]
    L.hipnegf_engine_sync.argtypes = [vp]
    L.hipnegf_engine_stream.argtypes = [vp]
    L.hipnegf_engine_stream.restype = vp
    L.hipnegf_struct_create.argtypes = [i, vp, i, vp, vp, i, i, C.POINTER(vp)]
    L.hipnegf_struct_create_surf.argtypes = [i, vp, i, vp, vp, vp, i, i, C.POINTER(vp)]
    L.hipnegf_struct_destroy.argtypes = [vp]
    L.hipnegf_struct_set_block.argtypes = [vp, i, i, i, vp]
    L.hipnegf_struct_set_block_device.argtypes = [vp, i, i, i, vp]
    L.hipnegf_solve_transmission.argtypes = [vp, vp, i, vp, vp, C.POINTER(Opts), vp, vp, vp]
    L.hipnegf_solve_transmission_device.argtypes = [vp, vp, i, vp, vp, C.POINTER(Opts), vp, vp, vp]
    L.hipnegf_solve_transmission_pairs.argtypes = [vp, vp, i, vp, vp, C.POINTER(Opts), i, vp, vp,
                                                   vp, vp, vp]
    L.hipnegf_solve_transmission_dos.argtypes = [vp, vp, i, vp, vp, C.POINTER(Opts), i, vp, vp,
                                                 C.POINTER(DosOpts), vp, vp, vp, vp]
    L.hipnegf_struct_set_tun_proj.argtypes = [vp, i, vp]
    L.hipnegf_density_layout.argtypes = [vp, vp]
    L.hipnegf_density_accumulate.argtypes = [vp, vp, i, vp, vp, vp, C.POINTER(Opts), i, vp, vp, i, i,
                                             vp, vp]
    L.hipnegf_solve_scba_mw.argtypes = [vp, vp, i, vp, vp, C.POINTER(Opts), vp, i, d, vp, i, i,
                                        vp, vp, vp, vp, vp, vp]
    L.hipnegf_zgemm_batched.argtypes = [vp, i, i, i, vp, vp, ll, i, i, vp, ll, i, i, vp, vp, ll, i, i]
    L.hipnegf_cgemm_batched.argtypes = L.hipnegf_zgemm_batched.argtypes
    L.hipnegf_zinverse_batched.argtypes = [vp, i, vp, ll, vp, ll, i, vp]
    L.hipnegf_cinverse_batched.argtypes = L.hipnegf_zinverse_batched.argtypes
    # block-op ABI
    L.hipnegf_blasInit.argtypes = [C.POINTER(vp)]
    L.hipnegf_blasFinalize.argtypes = [vp]
    L.hipnegf_ZmultMat.argtypes = [vp, C.c_size_t, C.c_size_t, C.c_size_t, vp, vp, vp, vp, vp, i]
    L.hipnegf_CmultMat.argtypes = L.hipnegf_ZmultMat.argtypes
    L.hipnegf_Zinverse.argtypes = [vp, vp, vp, vp, C.c_size_t]
    L.hipnegf_Cinverse.argtypes = L.hipnegf_Zinverse.argtypes
    L.hipnegf_Zinverse_noalloc.argtypes = [vp, vp, vp, C.c_size_t]
    L.hipnegf_createMat.argtypes = [C.POINTER(vp), C.c_size_t]
    L.hipnegf_deleteMat.argtypes = [C.POINTER(vp)]
    L.hipnegf_mallocAsync.argtypes = [C.POINTER(vp), C.c_size_t]
    L.hipnegf_freeAsync.argtypes = [C.POINTER(vp)]
    L.hipnegf_copyMatH2D.argtypes = [vp, vp, C.c_size_t]
    L.hipnegf_copyMatH2D_async.argtypes = [vp, vp, C.c_size_t]
    L.hipnegf_copyMatD2H.argtypes = [vp, vp, C.c_size_t]
    L.hipnegf_meminfo.argtypes = [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]
    L.hipnegf_getDeviceCount.argtypes = [C.POINTER(i)]
    L.hipnegf_getDevice.argtypes = [C.POINTER(i)]
    L.hipnegf_setDevice.argtypes = [i]
    L.hipnegf_solverInit.argtypes = [C.POINTER(vp)]
    L.hipnegf_solverFinalize.argtypes = [vp]
    L.hipnegf_Cmatcopy.argtypes = [vp, vp, vp, C.c_size_t]
    L.hipnegf_Zmatsum.argtypes = [vp, C.c_size_t, C.c_size_t, vp, vp, vp, vp, vp, i]
    L.hipnegf_Cmatsum.argtypes = L.hipnegf_Zmatsum.argtypes
    L.hipnegf_Zinitmat.argtypes = [vp, C.c_size_t]
    L.hipnegf_Cinitmat.argtypes = [vp, C.c_size_t]
    L.hipnegf_Ztrace.argtypes = [vp, vp, C.c_size_t, vp, i]
    L.hipnegf_Ztrace.restype = C.c_double
    L.hipnegf_Ctrace.argtypes = [vp, vp, C.c_size_t, vp, i]
    L.hipnegf_Ctrace.restype = C.c_float
    L.hipnegf_Zmatcopy.argtypes = [vp, vp, vp, C.c_size_t]
    L.hipnegf_Zasum.argtypes = [vp, vp, C.POINTER(C.c_double), C.c_size_t]
    L.hipnegf_Casum.argtypes = [vp, vp, C.POINTER(C.c_float), C.c_size_t]
    L.hipnegf_Zdecimation.argtypes = [vp, vp, vp, vp, vp, vp, C.c_size_t, i, C.POINTER(C.c_int), d]
    L.hipnegf_Cdecimation.argtypes = [vp, vp, vp, vp, vp, vp, C.c_size_t, i, C.POINTER(C.c_int),
                                      C.c_float]
    L.hipnegf_engine_set_sgf_cache.argtypes = [vp, ll]
    L.hipnegf_engine_sgf_stats.argtypes = [vp, C.POINTER(ll), C.POINTER(ll)]
    L.hipnegf_struct_set_fictitious.argtypes = [vp, i, i, d]
    L.hipnegf_profile_enable.argtypes = [i]
    L.hipnegf_calib_mfma.argtypes = [i, i, i, C.POINTER(d)]
    L.hipnegf_profile_read.argtypes = [i, C.POINTER(d), C.POINTER(d), C.POINTER(d),
                                       C.POINTER(ll)]
    _LIB = L
    return L


def check(rc: int):
    if rc != 0:
        msg = load().hipnegf_last_error().decode()
        raise HipNegfError(msg or f"hipnegf call failed (rc={rc})")
    return rc


# launch-timing tags (hipnegf_profile_*)
PT_GEMM_MASKED, PT_GEMM_SWEEP, PT_GEMM_INV, PT_INV_PANEL, PT_INV_SMALL, PT_FUSED = range(6)
# Kernel serving each tag for FP64 blocks with K, N >= 128 (the 64 x 128 tiling);
# smaller GEMMs of a tag run on k_gemm<double, tag>.
PT_NAMES = {PT_GEMM_MASKED: "k_gemm_w<0>", PT_GEMM_SWEEP: "k_gemm_w<1>",
            PT_GEMM_INV: "k_gemm_w<2>", PT_INV_PANEL: "k_inv_gj",
            PT_INV_SMALL: "k_lu_inv_2d", PT_FUSED: "k_t_fused"}


def gemm_force_variant(v: int):
    """FP64 GEMM tile: -1 autotuned, 0 always 64 x 64, 1 always 64 x 128, 2 always 32 x 32."""
    L = load()
    L.hipnegf_gemm_force_variant.argtypes = [C.c_int]
    check(L.hipnegf_gemm_force_variant(int(v)))


def gemm_tune_stats() -> dict:
    out = np.zeros(4, dtype=np.int64)
    L = load()
    L.hipnegf_gemm_tune_stats.argtypes = [C.c_void_p]
    check(L.hipnegf_gemm_tune_stats(_ptr(out)))
    return dict(shapes=int(out[0]), wide=int(out[1]), on=bool(out[2]), single_wave=int(out[3]))


def profile_enable(on: bool):
    check(load().hipnegf_profile_enable(1 if on else 0))


def profile_clear():
    check(load().hipnegf_profile_clear())


def profile_read(tag: int) -> dict:
    ms, fl, by, n = C.c_double(), C.c_double(), C.c_double(), C.c_longlong()
    check(load().hipnegf_profile_read(tag, C.byref(ms), C.byref(fl), C.byref(by), C.byref(n)))
    return dict(ms=ms.value, flops=fl.value, bytes=by.value, launches=n.value)


def profile_kernels() -> dict:
    """Launch statistics per kernel actually run: the GEMM tags split by the
    tile that served them (the autotuner may pick either FP64 tile)."""
    L = load()
    L.hipnegf_profile_read_variant.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_void_p,
                                               C.c_void_p, C.c_void_p]
    out = {}
    for t, name in PT_NAMES.items():
        if t in (PT_GEMM_MASKED, PT_GEMM_SWEEP, PT_GEMM_INV):
            for v, kname in ((1, f"k_gemm_w<{t}>"), (0, f"k_gemm<double, {t}>"), (2, f"k_gemm<float, {t}>"),
                             (3, f"k_gemm_s<{t}>")):
                ms, fl, by, n = C.c_double(), C.c_double(), C.c_double(), C.c_longlong()
                check(L.hipnegf_profile_read_variant(t, v, C.byref(ms), C.byref(fl), C.byref(by),
                                                     C.byref(n)))
                if n.value or v == 1:
                    out[kname] = dict(ms=ms.value, flops=fl.value, bytes=by.value, launches=n.value)
        else:
            out[name] = profile_read(t)
    return out


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


class Engine:
    """One engine = one HIP stream + a device workspace arena."""

    def __init__(self, device: int = 0):
        L = load()
        self._h = C.c_void_p()
        check(L.hipnegf_engine_create(device, C.byref(self._h)))

    @property
    def handle(self):
        return self._h

    def set_sgf_cache(self, nbytes: int):
        check(load().hipnegf_engine_set_sgf_cache(self._h, int(nbytes)))

    def sgf_stats(self):
        h, s = C.c_longlong(), C.c_longlong()
        check(load().hipnegf_engine_sgf_stats(self._h, C.byref(h), C.byref(s)))
        return h.value, s.value

    def sync(self):
        check(load().hipnegf_engine_sync(self._h))

    def close(self):
        if self._h:
            load().hipnegf_engine_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _opts(delta: float, ni: int, nf: int, precision: int, max_batch: int, ns_steps=0,
          polish=0, eps_drop=0.0, flags=0, lanes=1) -> Opts:
    return Opts(delta, ni, nf, precision, max_batch, ns_steps, polish, eps_drop, flags, lanes)


class DeviceStructure:
    """Device copy of one or more BlockStructures of identical shape (slots)."""

    def __init__(self, structs: Sequence[BlockStructure] | BlockStructure, nslots: Optional[int] = None):
        if isinstance(structs, BlockStructure):
            structs = [structs]
        s0 = structs[0]
        self.template = s0
        self.nslots = nslots or len(structs)
        smode = 0
        if s0.Sd is not None:
            smode |= S_DIAG
        if s0.Su is not None or s0.Sl is not None:
            smode |= S_OFFDIAG
        if any(c.S00 is not None for c in s0.contacts):
            smode |= S_CONTACT
        self.smode = smode
        bsz = np.array(s0.bsz, dtype=np.int32)
        cblk = np.array([c.cblk for c in s0.contacts], dtype=np.int32)
        nc = np.array([c.nc for c in s0.contacts], dtype=np.int32)
        ns = np.array([c.ns for c in s0.contacts], dtype=np.int32)
        self._h = C.c_void_p()
        if ns.any():
            check(load().hipnegf_struct_create_surf(s0.nbl, _ptr(bsz), len(cblk), _ptr(cblk), _ptr(nc),
                                                    _ptr(ns), smode, self.nslots, C.byref(self._h)))
        else:
            check(load().hipnegf_struct_create(s0.nbl, _ptr(bsz), len(cblk), _ptr(cblk), _ptr(nc),
                                               smode, self.nslots, C.byref(self._h)))
        for k, s in enumerate(structs):
            self.set_slot(k, s)

    @classmethod
    def empty(cls, n: int, nbl: int, cblk=(None, 1), smode: int = 0, nslots: int = 1):
        """Zero-filled structure of nbl PLs of size n with contacts of size n on
        blocks ``cblk`` (to be filled on the device, e.g. by a generator)."""
        from types import SimpleNamespace
        self = cls.__new__(cls)
        cb = [nbl if c is None else c for c in cblk]
        self.template = SimpleNamespace(ncont=len(cb), nbl=nbl, bsz=[n] * nbl, ndev=n * nbl)
        self.nslots = nslots
        self.smode = smode
        bsz = np.full(nbl, n, dtype=np.int32)
        cba = np.array(cb, dtype=np.int32)
        nc = np.full(len(cb), n, dtype=np.int32)
        self._h = C.c_void_p()
        check(load().hipnegf_struct_create(nbl, _ptr(bsz), len(cba), _ptr(cba), _ptr(nc), smode,
                                           nslots, C.byref(self._h)))
        return self

    def set_tun_proj(self, idx=None):
        """tun_proj: 1-based device orbitals whose rows enter the transmission
        traces (None or empty: all)."""
        a = np.ascontiguousarray([] if idx is None else idx, dtype=np.int32)
        check(load().hipnegf_struct_set_tun_proj(self._h, len(a), _ptr(a) if len(a) else None))

    def set_fictitious(self, contact: int, dos: float, fictitious: bool = True):
        """contact 0-based; g_s = -i pi dos I for that contact."""
        check(load().hipnegf_struct_set_fictitious(self._h, contact, 1 if fictitious else 0, dos))

    def set_block(self, slot: int, kind: int, index: int, a: np.ndarray):
        a = np.asfortranarray(a, dtype=np.complex128)
        check(load().hipnegf_struct_set_block(self._h, slot, kind, index, _ptr(a)))

    def set_slot(self, slot: int, s: BlockStructure):
        for i, a in enumerate(s.Hd):
            self.set_block(slot, HD, i, a)
        for i in range(s.nbl - 1):
            self.set_block(slot, HU, i, s.Hu[i])
            self.set_block(slot, HL, i, s.Hl[i])
        if self.smode & S_DIAG:
            for i, a in enumerate(s.Sd):
                self.set_block(slot, SD, i, a)
        if self.smode & S_OFFDIAG:
            for i in range(s.nbl - 1):
                self.set_block(slot, SU, i, s.Su[i] if s.Su is not None else np.zeros_like(s.Hu[i]))
                self.set_block(slot, SL, i, s.Sl[i] if s.Sl is not None else np.zeros_like(s.Hl[i]))
        for c, ct in enumerate(s.contacts):
            self.set_block(slot, H00, c, ct.H00)
            self.set_block(slot, H01, c, ct.H01)
            self.set_block(slot, HMC, c, ct.HMC)
            if ct.ns:
                self.set_block(slot, HSS, c, ct.Hss)
                self.set_block(slot, HS1, c, ct.Hs1)
                self.set_block(slot, H1S, c, ct.H1s)
            if self.smode & S_CONTACT:
                n, q = ct.nc, ct.ns
                self.set_block(slot, S00, c, ct.S00 if ct.S00 is not None else np.eye(n))
                self.set_block(slot, S01, c, ct.S01 if ct.S01 is not None else np.zeros((n, n)))
                self.set_block(slot, SMC, c, ct.SMC if ct.SMC is not None else np.zeros_like(ct.HMC))
                if q:
                    self.set_block(slot, SSS, c, ct.Sss if ct.Sss is not None else np.eye(q))
                    self.set_block(slot, SS1, c, ct.Ss1 if ct.Ss1 is not None else np.zeros((q, n)))
                    self.set_block(slot, S1S, c, ct.S1s if ct.S1s is not None else np.zeros((n, q)))

    @property
    def handle(self):
        return self._h

    def close(self):
        if self._h:
            load().hipnegf_struct_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def solve_transmission(eng: Engine, ds: DeviceStructure, E: np.ndarray, slot=None,
                       delta: float = 1e-4, ni: int = 1, nf: int = 2, precision: int = FP64,
                       max_batch: int = 0, flags: int = 0, lanes: int = 1, ns_steps: int = 0):
    """Host-array batched transmission.  Returns (T, ncyc[n, ncont], status).
    precision=MIXED: FP32 sweep, ns_steps FP64 Newton-Schulz steps per block
    inverse (0: default 1, -1: none)."""
    E = np.ascontiguousarray(E, dtype=np.float64)
    n = len(E)
    sl = np.zeros(n, dtype=np.int32) if slot is None else np.ascontiguousarray(slot, dtype=np.int32)
    ncont = ds.template.ncont
    T = np.zeros(n)
    ncyc = np.zeros((n, max(ncont, 1)), dtype=np.int32)
    st = np.zeros(n, dtype=np.int32)
    o = _opts(delta, ni, nf, precision, max_batch, ns_steps=ns_steps, flags=flags, lanes=lanes)
    check(load().hipnegf_solve_transmission(eng.handle, ds.handle, n, _ptr(E), _ptr(sl),
                                            C.byref(o), _ptr(T), _ptr(ncyc), _ptr(st)))
    return T, ncyc, st


def solve_transmission_pairs(eng: Engine, ds: DeviceStructure, E: np.ndarray, pairs, slot=None,
                             delta: float = 1e-4, precision: int = FP64, max_batch: int = 0,
                             flags: int = 0):
    """T for several (ni, nf) contact pairs (1-based) per energy: returns
    (T[n, npairs], ncyc[n, ncont], status[n])."""
    E = np.ascontiguousarray(E, dtype=np.float64)
    n = len(E)
    sl = np.zeros(n, dtype=np.int32) if slot is None else np.ascontiguousarray(slot, dtype=np.int32)
    ni = np.ascontiguousarray([p[0] for p in pairs], dtype=np.int32)
    nf = np.ascontiguousarray([p[1] for p in pairs], dtype=np.int32)
    T = np.zeros((n, len(pairs)))
    nc = max(ds.template.ncont, 1)
    ncyc = np.zeros((n, nc), dtype=np.int32)
    st = np.zeros(n, dtype=np.int32)
    o = _opts(delta, 1, 2, precision, max_batch, flags=flags)
    check(load().hipnegf_solve_transmission_pairs(eng.handle, ds.handle, n, _ptr(E), _ptr(sl),
                                                  C.byref(o), len(pairs), _ptr(ni), _ptr(nf),
                                                  _ptr(T), _ptr(ncyc), _ptr(st)))
    return T, ncyc, st


class DosOpts(C.Structure):
    _fields_ = [("dos_delta", C.c_double), ("nldos", C.c_int), ("full", C.c_int),
                ("ptr", C.c_void_p), ("idx", C.c_void_p)]


def solve_transmission_dos(eng: Engine, ds: DeviceStructure, E: np.ndarray, pairs=((1, 2),),
                           proj=None, full: bool = False, dos_delta: float = 0.0, slot=None,
                           delta: float = 1e-4, precision: int = FP64, max_batch: int = 0,
                           flags: int = 0, lanes: int = 1):
    """Transmission and local DOS per energy (calculate_transmissions_and_dos).
    proj: list of 1-based orbital lists (LDOS projections); full: one column
    per device orbital.  Returns (T[n, npairs], ldos[n, nout], ncyc[n, ncont],
    status[n])."""
    E = np.ascontiguousarray(E, dtype=np.float64)
    n = len(E)
    sl = np.zeros(n, dtype=np.int32) if slot is None else np.ascontiguousarray(slot, dtype=np.int32)
    ni = np.ascontiguousarray([p[0] for p in pairs], dtype=np.int32)
    nf = np.ascontiguousarray([p[1] for p in pairs], dtype=np.int32)
    lists = [] if (proj is None or full) else list(proj)
    ptr = np.zeros(len(lists) + 1, dtype=np.int32)
    for k, l in enumerate(lists):
        ptr[k + 1] = ptr[k] + len(l)
    idx = np.ascontiguousarray(np.concatenate([np.asarray(l, dtype=np.int32) for l in lists])
                               if lists else np.zeros(1, np.int32), dtype=np.int32)
    nout = ds.template.ndev if full else len(lists)
    d = DosOpts(dos_delta, len(lists), 1 if full else 0, ptr.ctypes.data, idx.ctypes.data)
    T = np.zeros((n, len(pairs)))
    ldos = np.zeros((n, max(nout, 1)))
    nc = max(ds.template.ncont, 1)
    ncyc = np.zeros((n, nc), dtype=np.int32)
    st = np.zeros(n, dtype=np.int32)
    o = _opts(delta, 1, 2, precision, max_batch, flags=flags, lanes=lanes)
    check(load().hipnegf_solve_transmission_dos(eng.handle, ds.handle, n, _ptr(E), _ptr(sl), C.byref(o),
                                                len(pairs), _ptr(ni), _ptr(nf), C.byref(d), _ptr(T),
                                                _ptr(ldos), _ptr(ncyc), _ptr(st)))
    return T, ldos[:, :nout], ncyc, st


def density_layout(ds: DeviceStructure) -> np.ndarray:
    t = ds.template
    off = np.zeros(3 * t.nbl - 2 + 2 * t.ncont + 1, dtype=np.int64)
    check(load().hipnegf_density_layout(ds.handle, _ptr(off)))
    return off


def density_blocks(ds: DeviceStructure, acc: np.ndarray) -> dict:
    """Unpack a density accumulator into dict(Gd, Gu, Gl, Glc, Gcl) of arrays."""
    t = ds.template
    off = density_layout(ds)
    b = t.bsz
    nd = [b[c.cblk - 1] for c in t.contacts]
    nc = [c.nc + c.ns for c in t.contacts]  # outer blocks span surface + PL1
    shapes = ([(n, n) for n in b] + [(b[i], b[i + 1]) for i in range(t.nbl - 1)] +
              [(b[i + 1], b[i]) for i in range(t.nbl - 1)] + [(nd[k], nc[k]) for k in range(t.ncont)] +
              [(nc[k], nd[k]) for k in range(t.ncont)])
    mats = [acc[off[j]:off[j + 1]].reshape(sh, order="F") for j, sh in enumerate(shapes)]
    n1, n2 = t.nbl, t.nbl - 1
    return dict(Gd=mats[:n1], Gu=mats[n1:n1 + n2], Gl=mats[n1 + n2:n1 + 2 * n2],
                Glc=mats[n1 + 2 * n2:n1 + 2 * n2 + t.ncont], Gcl=mats[n1 + 2 * n2 + t.ncont:])


def density_accumulate(eng: Engine, ds: DeviceStructure, z: np.ndarray, w: np.ndarray, mode: int = 0,
                       frm=None, ref: int = 0, outer: int = 2, slot=None, delta: float = 1e-4,
                       acc: Optional[np.ndarray] = None, max_batch: int = 0):
    """Sum of w_i G^r(z_i) (mode 0, contour points) or w_i G^n(Re z_i)
    (mode 1, real axis; contacts at Re z + i delta; frm[n, ncont+1]) over
    the blocks of the structure.  Returns (acc, status)."""
    z = np.ascontiguousarray(z, dtype=np.complex128)
    n = len(z)
    Er = np.ascontiguousarray(z.real)
    Ei = np.ascontiguousarray(z.imag)
    ww = np.ascontiguousarray(w, dtype=np.complex128)
    sl = np.zeros(n, dtype=np.int32) if slot is None else np.ascontiguousarray(slot, dtype=np.int32)
    tot = int(density_layout(ds)[-1])
    a = np.zeros(tot, dtype=np.complex128) if acc is None else np.ascontiguousarray(acc, dtype=np.complex128)
    fr = None if frm is None else np.ascontiguousarray(frm, dtype=np.float64)
    st = np.zeros(n, dtype=np.int32)
    o = _opts(delta, 1, 2, FP64, max_batch)
    check(load().hipnegf_density_accumulate(eng.handle, ds.handle, n, _ptr(Er), _ptr(Ei), _ptr(sl),
                                            C.byref(o), mode, _ptr(ww), None if fr is None else _ptr(fr),
                                            ref, outer, _ptr(a), _ptr(st)))
    return a, st


def solve_scba_mw(eng: Engine, ds: DeviceStructure, E: np.ndarray, coupling: np.ndarray,
                  niter: int, tol: float, frm, ref: int, slot=None, delta: float = 1e-4,
                  precision: int = FP64, max_batch: int = 0, want_diag: bool = False, lanes: int = 1):
    """Elastic-dephasing SCBA + Meir-Wingreen per energy (host arrays).
    frm: [ncont+1] shared or [n, ncont+1] per item.  Returns dict."""
    E = np.ascontiguousarray(E, dtype=np.float64)
    n = len(E)
    s = ds.template
    sl = np.zeros(n, dtype=np.int32) if slot is None else np.ascontiguousarray(slot, dtype=np.int32)
    fr = np.ascontiguousarray(frm, dtype=np.float64)
    per_item = 1 if fr.ndim == 2 else 0
    cp = np.ascontiguousarray(coupling, dtype=np.float64)
    curr = np.zeros((n, s.ncont))
    iters = np.zeros(n, dtype=np.int32)
    err = np.zeros(n)
    st = np.zeros(n, dtype=np.int32)
    gn = np.zeros((n, s.ndev), dtype=np.complex128) if want_diag else None
    gr = np.zeros((n, s.ndev), dtype=np.complex128) if want_diag else None
    o = _opts(delta, 1, 2, precision, max_batch, lanes=lanes)
    check(load().hipnegf_solve_scba_mw(
        eng.handle, ds.handle, n, _ptr(E), _ptr(sl), C.byref(o), _ptr(cp), niter, tol, _ptr(fr),
        per_item, ref, _ptr(curr), _ptr(iters), _ptr(err), _ptr(st),
        None if gn is None else _ptr(gn), None if gr is None else _ptr(gr)))
    return dict(curr=curr, iters=iters, err=err, status=st, gn_diag=gn, gr_diag=gr)


def inel_kmat_po(ndz: int, dz: float, kfrac, kweights, coupling: float, wq: float, cell_area: float,
                 eps0: float, eps_inf: float, q0: float, latt=None) -> np.ndarray:
    """K table of the polar-optical model (prepare_POKmat): (ndz, nk, nk), Fortran order."""
    kf = np.ascontiguousarray(np.asarray(kfrac, dtype=np.float64).reshape(-1, 3))
    nk = kf.shape[0]
    w = np.ascontiguousarray(kweights, dtype=np.float64)
    la = np.ascontiguousarray(np.zeros(9) if latt is None else np.asarray(latt, float).T.ravel())
    K = np.zeros(ndz * nk * nk)
    L = load()
    L.hipnegf_inel_kmat_po.argtypes = [C.c_int, C.c_double, C.c_int, C.c_void_p, C.c_void_p,
                                       C.c_void_p] + [C.c_double] * 6 + [C.c_void_p]
    check(L.hipnegf_inel_kmat_po(ndz, dz, nk, _ptr(kf), _ptr(w), _ptr(la), coupling, wq, cell_area,
                                 eps0, eps_inf, q0, _ptr(K)))
    return K.reshape((ndz, nk, nk), order="F")


def inel_kmat_nonpo(ndz: int, dz: float, kfrac, kweights, coupling: float, D0: float,
                    latt=None) -> np.ndarray:
    """K table of the non-polar optical model (prepare_NonPOKmat)."""
    kf = np.ascontiguousarray(np.asarray(kfrac, dtype=np.float64).reshape(-1, 3))
    nk = kf.shape[0]
    w = np.ascontiguousarray(kweights, dtype=np.float64)
    la = np.ascontiguousarray(np.zeros(9) if latt is None else np.asarray(latt, float).T.ravel())
    K = np.zeros(ndz * nk * nk)
    L = load()
    L.hipnegf_inel_kmat_nonpo.argtypes = [C.c_int, C.c_double, C.c_int, C.c_void_p, C.c_void_p,
                                          C.c_void_p, C.c_double, C.c_double, C.c_void_p]
    check(L.hipnegf_inel_kmat_nonpo(ndz, dz, nk, _ptr(kf), _ptr(w), _ptr(la), coupling, D0, _ptr(K)))
    return K.reshape((ndz, nk, nk), order="F")


def solve_inelastic_mw(eng: Engine, ds: DeviceStructure, E: np.ndarray, kweights, frm, ref: int,
                       K: np.ndarray, iz, shift: int, Nq: float, tridiag: bool, niter: int,
                       tol: float, coupling=None, niter_ela: int = 0, tol_ela: float = 1e-7,
                       integration: int = 0, g_spin: float = 2.0, delta: float = 1e-4,
                       max_batch: int = 0, lanes: int = 1) -> dict:
    """meir_wingreen with an inelastic model over the (E,k) grid: E uniform,
    the first len(kweights) slots of ds as the k-points, frm[ne, ncont+1].
    Returns dict(curr_mat[ne, ncont], currents, passes, err, status[nk, ne],
    iters[nk, ne])."""
    E = np.ascontiguousarray(E, dtype=np.float64)
    ne = len(E)
    w = np.ascontiguousarray(kweights, dtype=np.float64)
    nk = len(w)
    ncont = ds.template.ncont
    fr = np.ascontiguousarray(np.asarray(frm, dtype=np.float64).reshape(ne, ncont + 1))
    Kc = np.ascontiguousarray(np.asarray(K, dtype=np.float64).ravel(order="F"))
    ndz = np.asarray(K).shape[0]
    izc = np.ascontiguousarray(iz, dtype=np.int32)
    cp = None if coupling is None else np.ascontiguousarray(coupling, dtype=np.float64)
    cm = np.zeros(ne * ncont)
    cur = np.zeros(ncont)
    npass = C.c_int()
    err = C.c_double()
    st = np.zeros(ne * nk, dtype=np.int32)
    it = np.zeros(ne * nk, dtype=np.int32)
    o = _opts(delta, 1, 2, FP64, max_batch, lanes=lanes)
    L = load()
    vp, i, d = C.c_void_p, C.c_int, C.c_double
    L.hipnegf_solve_inelastic_mw.argtypes = [vp, vp, i, vp, i, vp, C.POINTER(Opts), vp, i, d, vp, i, vp,
                                             i, vp, i, d, i, i, d, i, d, vp, vp, C.POINTER(C.c_int),
                                             C.POINTER(C.c_double), vp, vp]
    check(L.hipnegf_solve_inelastic_mw(
        eng.handle, ds.handle, ne, _ptr(E), nk, _ptr(w), C.byref(o), None if cp is None else _ptr(cp),
        niter_ela, tol_ela, _ptr(fr), ref, _ptr(Kc), ndz, _ptr(izc), shift, Nq, 1 if tridiag else 0,
        niter, tol, integration, g_spin, _ptr(cm), _ptr(cur), C.byref(npass), C.byref(err), _ptr(st),
        _ptr(it)))
    return dict(curr_mat=cm.reshape((ne, ncont), order="F"), currents=cur, passes=npass.value,
                err=err.value, status=st.reshape(nk, ne), iters=it.reshape(nk, ne))
