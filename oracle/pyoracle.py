"""ctypes front end of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from typing import List, Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

zptr = C.POINTER(C.c_double)  # complex arrays are passed as interleaved doubles


def build(force: bool = False) -> str:
    so = os.path.join(_HERE, "liboracle.so")
    if force or not os.path.exists(so):
        subprocess.check_call(["make", "-C", _HERE, "-B" if force else "all"],
                              stdout=subprocess.DEVNULL)
    return so


def lib():
    global _LIB
    if _LIB is None:
        so = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(so):
            build()
        L = C.CDLL(so)
        L.or_set_backend.argtypes = [C.c_int]
        L.or_backend_name.restype = C.c_char_p
        L.or_transmission.restype = C.c_double
        L.or_transmission.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_int, C.c_int,
                                      C.c_void_p, C.c_double, C.c_double]
        L.or_transmission_sigma.restype = C.c_double
        L.or_transmission_sigma.argtypes = [C.c_void_p, C.c_double, C.c_void_p, C.c_int, C.c_int,
                                            C.c_double]
        L.or_transmission_sigma_tun.restype = C.c_double
        L.or_transmission_sigma_tun.argtypes = [C.c_void_p, C.c_double, C.c_void_p, C.c_int, C.c_int,
                                                C.c_double, C.c_int, C.c_void_p]
        L.or_transmission_dos.restype = C.c_int
        L.or_transmission_dos.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_int,
                                          C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_int,
                                          C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p,
                                          C.c_void_p, C.c_void_p, C.c_double, C.c_double]
        L.or_transmission_dos_sigma.restype = C.c_int
        L.or_transmission_dos_sigma.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_void_p,
                                                C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p,
                                                C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p,
                                                C.c_void_p, C.c_void_p, C.c_double]
        L.or_transmission_batch.restype = C.c_double
        L.or_transmission_batch.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_double, C.c_int,
                                            C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_double]
        L.or_contact_sigma.restype = C.c_int
        L.or_contact_sigma.argtypes = [C.c_double, C.c_double, C.c_int, C.c_int] + [C.c_void_p] * 8 + \
            [C.c_void_p, C.c_double]
        L.or_contact_sigma_surf.restype = C.c_int
        L.or_contact_sigma_surf.argtypes = [C.c_double, C.c_double, C.c_int, C.c_int, C.c_int] + \
            [C.c_void_p] * 14 + [C.c_void_p, C.c_double]
        L.or_decimation.restype = C.c_int
        L.or_decimation.argtypes = [C.c_int] + [C.c_void_p] * 5 + [C.c_double]
        L.or_zinv.restype = C.c_int
        L.or_zinv.argtypes = [C.c_int, C.c_void_p, C.c_void_p]
        L.or_scba_meir_wingreen.restype = C.c_int
        L.or_scba_meir_wingreen.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_void_p, C.c_int,
                                            C.c_double, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p,
                                            C.c_void_p, C.c_void_p, C.c_double, C.c_double]
        L.or_scba_mw_sigma.restype = C.c_int
        L.or_scba_mw_sigma.argtypes = [C.c_void_p, C.c_double, C.c_void_p, C.c_void_p, C.c_int,
                                       C.c_double, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p,
                                       C.c_void_p, C.c_void_p, C.c_double]
        L.or_gr_tridiag.restype = C.c_int
        L.or_gr_tridiag.argtypes = [C.c_void_p, C.c_double, C.c_void_p, C.c_void_p, C.c_void_p,
                                    C.c_void_p]
        L.or_fermi.restype = C.c_double
        L.or_fermi.argtypes = [C.c_double] * 3
        L.or_integrate_el.restype = C.c_double
        L.or_integrate_el.argtypes = [C.c_void_p, C.c_int] + [C.c_double] * 7
        L.or_integrate_el_mw.restype = C.c_double
        L.or_integrate_el_mw.argtypes = [C.c_int, C.c_void_p, C.c_int, C.c_double]
        L.or_npoints.restype = C.c_int
        L.or_npoints.argtypes = [C.c_double] * 3
        L.or_omp_max_threads.restype = C.c_int
        L.or_gauleg.argtypes = [C.c_double, C.c_double, C.c_void_p, C.c_void_p, C.c_int]
        L.or_contour_grid.restype = C.c_int
        L.or_contour_grid.argtypes = [C.c_double, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int,
                                      C.c_int, C.c_double, C.c_void_p, C.c_void_p]
        L.or_efa_contour_grid.restype = C.c_int
        L.or_efa_contour_grid.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.c_int, C.c_int,
                                          C.c_int, C.c_int, C.c_double, C.c_void_p, C.c_void_p]
        L.or_trapez.argtypes = [C.c_double, C.c_double, C.c_void_p, C.c_void_p, C.c_int]
        L.or_real_axis_grid.restype = C.c_int
        L.or_real_axis_grid.argtypes = [C.c_double, C.c_double, C.c_double, C.c_int, C.c_int,
                                        C.c_double, C.c_double, C.c_void_p, C.c_void_p]
        L.or_gr_point.restype = C.c_int
        L.or_gr_point.argtypes = [C.c_void_p, C.c_double, C.c_double] + [C.c_void_p] * 6 + [C.c_double]
        L.or_gr_point_real.restype = C.c_int
        L.or_gr_point_real.argtypes = [C.c_void_p, C.c_double, C.c_double] + [C.c_void_p] * 6 + [C.c_double]
        L.or_gn_point.restype = C.c_int
        L.or_gn_point.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_void_p, C.c_int] + \
            [C.c_void_p] * 6 + [C.c_double, C.c_double]
        L.dd_transmission_batch.restype = C.c_int
        L.dd_transmission_batch.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_double,
                                            C.c_double, C.c_void_p, C.c_void_p, C.c_int]
        _LIB = L
    return _LIB


def set_backend(which: str) -> str:
    """'builtin', 'mkl', or 'gj' (built-in products with the Gauss-Jordan
    inverse of the GPU's blocked path, for accuracy experiments); returns the
    backend name actually in use."""
    r = lib().or_set_backend({"mkl": 1, "gj": 2}.get(which, 0))
    if r < 0:
        raise RuntimeError("MKL backend requested but libmkl_rt.so could not be loaded")
    return lib().or_backend_name().decode()


class _OrStruct(C.Structure):
    _fields_ = [("nbl", C.c_int), ("bsz", C.c_void_p), ("Hd", C.c_void_p), ("Sd", C.c_void_p),
                ("Hu", C.c_void_p), ("Su", C.c_void_p), ("Hl", C.c_void_p), ("Sl", C.c_void_p),
                ("ncont", C.c_int), ("cblk", C.c_void_p), ("nc", C.c_void_p),
                ("H00", C.c_void_p), ("S00", C.c_void_p), ("H01", C.c_void_p),
                ("S01", C.c_void_p), ("HMC", C.c_void_p), ("SMC", C.c_void_p),
                ("ns", C.c_void_p), ("Hss", C.c_void_p), ("Sss", C.c_void_p), ("Hs1", C.c_void_p),
                ("Ss1", C.c_void_p), ("H1s", C.c_void_p), ("S1s", C.c_void_p)]


def _cz(a: np.ndarray) -> np.ndarray:
    return np.asfortranarray(a, dtype=np.complex128)


class OracleStruct:
    """Keeps the numpy buffers alive while the C struct points at them."""

    def __init__(self, s):
        keep = []

        def parr(mats: Optional[Sequence[Optional[np.ndarray]]], n: int):
            if mats is None:
                return None
            arr = (C.c_void_p * max(n, 1))()
            for i, m in enumerate(mats):
                if m is None:
                    arr[i] = None
                else:
                    m = _cz(m)
                    keep.append(m)
                    arr[i] = m.ctypes.data
            keep.append(arr)
            return C.cast(arr, C.c_void_p)

        nbl = s.nbl
        bsz = np.array(s.bsz, dtype=np.int32)
        cblk = np.array([c.cblk for c in s.contacts], dtype=np.int32)
        nc = np.array([c.nc for c in s.contacts], dtype=np.int32)
        keep += [bsz, cblk, nc]
        st = _OrStruct()
        st.nbl = nbl
        st.bsz = bsz.ctypes.data
        st.Hd = parr(s.Hd, nbl)
        st.Sd = parr(s.Sd, nbl)
        st.Hu = parr(s.Hu, nbl - 1)
        st.Su = parr(s.Su, nbl - 1)
        st.Hl = parr(s.Hl, nbl - 1)
        st.Sl = parr(s.Sl, nbl - 1)
        st.ncont = len(s.contacts)
        st.cblk = cblk.ctypes.data
        st.nc = nc.ctypes.data
        cs = s.contacts
        st.H00 = parr([c.H00 for c in cs], len(cs))
        st.H01 = parr([c.H01 for c in cs], len(cs))
        st.HMC = parr([c.HMC for c in cs], len(cs))
        anyS = any(c.S00 is not None for c in cs)
        st.S00 = parr([c.S00 for c in cs], len(cs)) if anyS else None
        st.S01 = parr([c.S01 for c in cs], len(cs)) if anyS else None
        st.SMC = parr([c.SMC for c in cs], len(cs)) if anyS else None
        if any(c.ns for c in cs):
            ns = np.array([c.ns for c in cs], dtype=np.int32)
            keep.append(ns)
            st.ns = ns.ctypes.data
            for name in ("Hss", "Sss", "Hs1", "Ss1", "H1s", "S1s"):
                mats = [getattr(c, name) for c in cs]
                setattr(st, name, parr(mats, len(cs)) if any(m is not None for m in mats) else None)
        self._keep = keep
        self.c = st
        self.s = s

    @property
    def ptr(self):
        return C.byref(self.c)


SGFACC = 1e-9  # ln_precision.F90:40


def transmission(s, E: float, delta: float, ni: int = 1, nf: int = 2, eps_drop: float = 0.0):
    os_ = OracleStruct(s)
    ncyc = np.zeros(s.ncont, dtype=np.int32)
    T = lib().or_transmission(os_.ptr, E, delta, ni, nf, ncyc.ctypes.data, SGFACC, eps_drop)
    return T, ncyc


def transmission_batch(s, E: np.ndarray, delta: float, ni: int = 1, nf: int = 2,
                       nthreads: int = 0):
    os_ = OracleStruct(s)
    E = np.ascontiguousarray(E, dtype=np.float64)
    T = np.zeros(len(E))
    ncyc = np.zeros((len(E), s.ncont), dtype=np.int32)
    secs = lib().or_transmission_batch(os_.ptr, len(E), E.ctypes.data, delta, ni, nf,
                                       T.ctypes.data, ncyc.ctypes.data, nthreads, SGFACC)
    return T, ncyc, secs


def transmission_dd(s, E: np.ndarray, delta: float, nthreads: int = 0):
    """T(E) of the fast path in double-double arithmetic (negf_dd.c): returns
    (T_hi, T_lo, ncyc[npts, 2]); T_hi + T_lo is the value."""
    os_ = OracleStruct(s)
    E = np.ascontiguousarray(np.atleast_1d(E), dtype=np.float64)
    T = np.zeros((len(E), 2))
    ncyc = np.zeros((len(E), s.ncont), dtype=np.int32)
    r = lib().dd_transmission_batch(os_.ptr, len(E), E.ctypes.data, delta, SGFACC, T.ctypes.data,
                                    ncyc.ctypes.data, nthreads or (os.cpu_count() or 1))
    if r != 0:
        raise ValueError("dd_transmission: structure outside the fast-path scope or singular block")
    return T[:, 0].copy(), T[:, 1].copy(), ncyc


def contact_sigma(c, nd: int, z: complex, okco: Optional[np.ndarray] = None):
    """Returns (Sigma nd x nd, GS (ns+nc) x (ns+nc), ncyc)."""
    ng = c.nc + c.ns
    Sig = np.zeros((nd, nd), dtype=np.complex128, order="F")
    GS = np.zeros((ng, ng), dtype=np.complex128, order="F")
    opt = lambda a: None if a is None else _cz(a)
    bufs = [_cz(c.H00), opt(c.S00), _cz(c.H01), opt(c.S01), opt(c.Hss), opt(c.Sss), opt(c.Hs1),
            opt(c.Ss1), opt(c.H1s), opt(c.S1s), _cz(c.HMC), opt(c.SMC)]
    ptrs = [None if b is None else b.ctypes.data for b in bufs]
    ok = None if okco is None else okco.ctypes.data
    ncyc = lib().or_contact_sigma_surf(z.real, z.imag, c.nc, c.ns, nd, *ptrs, Sig.ctypes.data,
                                       GS.ctypes.data, ok, SGFACC)
    return Sig, GS, ncyc


def transmission_sigma(s, E: float, sigmas: List[np.ndarray], ni=1, nf=2, eps_drop=0.0):
    os_ = OracleStruct(s)
    sg = [_cz(x) for x in sigmas]
    arr = (C.c_void_p * len(sg))(*[x.ctypes.data for x in sg])
    return lib().or_transmission_sigma(os_.ptr, E, C.cast(arr, C.c_void_p), ni, nf, eps_drop)


def transmission_sigma_tun(s, E: float, sigmas: List[np.ndarray], ni=1, nf=2, eps_drop=0.0,
                           tun=None):
    os_ = OracleStruct(s)
    sg = [_cz(x) for x in sigmas]
    arr = (C.c_void_p * len(sg))(*[x.ctypes.data for x in sg])
    tv = np.ascontiguousarray(tun if tun is not None else [0], dtype=np.int32)
    return lib().or_transmission_sigma_tun(os_.ptr, E, C.cast(arr, C.c_void_p), ni, nf, eps_drop,
                                           0 if tun is None else len(tv), tv.ctypes.data)


def proj_csr(lists):
    """LDOS projections (list of 1-based orbital index lists) -> (ptr, idx) int32."""
    ptr = np.zeros(len(lists) + 1, dtype=np.int32)
    for i, l in enumerate(lists):
        ptr[i + 1] = ptr[i] + len(l)
    idx = np.concatenate([np.asarray(l, dtype=np.int32) for l in lists]) if lists else np.zeros(0, np.int32)
    return ptr, np.ascontiguousarray(idx, dtype=np.int32)


def transmission_dos(s, E: float, delta: float, dos_delta: float, pairs=((1, 2),), proj=None,
                     full: bool = False, tun=None, eps_drop: float = 0.0, sigmas=None):
    """calculate_transmissions_and_dos at one energy.  Returns dict(T[npairs],
    ldos[nldos or ndev], diag[ndev], ncyc[ncont])."""
    os_ = OracleStruct(s)
    ni = np.array([p[0] for p in pairs], dtype=np.int32)
    nf = np.array([p[1] for p in pairs], dtype=np.int32)
    lists = proj or []
    ptr, idx = proj_csr(lists)
    tv = np.ascontiguousarray(tun if tun is not None else [0], dtype=np.int32)
    ntun = 0 if tun is None else len(tv)
    T = np.zeros(len(pairs))
    ledos = np.zeros(s.ndev if full else max(len(lists), 1))
    diag = np.zeros(s.ndev)
    ncyc = np.zeros(max(s.ncont, 1), dtype=np.int32)
    nl = len(lists)
    if sigmas is None:
        r = lib().or_transmission_dos(os_.ptr, E, delta, dos_delta, len(pairs), ni.ctypes.data,
                                      nf.ctypes.data, ntun, tv.ctypes.data, nl, ptr.ctypes.data,
                                      idx.ctypes.data, 1 if full else 0, T.ctypes.data,
                                      ledos.ctypes.data, diag.ctypes.data, ncyc.ctypes.data, SGFACC,
                                      eps_drop)
    else:
        sg = [_cz(x) for x in sigmas]
        arr = (C.c_void_p * max(len(sg), 1))(*[x.ctypes.data for x in sg])
        r = lib().or_transmission_dos_sigma(os_.ptr, E, dos_delta, C.cast(arr, C.c_void_p), len(pairs),
                                            ni.ctypes.data, nf.ctypes.data, ntun, tv.ctypes.data, nl,
                                            ptr.ctypes.data, idx.ctypes.data, 1 if full else 0,
                                            T.ctypes.data, ledos.ctypes.data, diag.ctypes.data,
                                            eps_drop)
    if r != 0:
        raise ValueError("LDOS index < 1")
    return dict(T=T, ldos=ledos[:s.ndev if full else nl], diag=diag, ncyc=ncyc[:s.ncont])


def zinv(A: np.ndarray):
    A = _cz(A)
    out = np.zeros_like(A)
    info = lib().or_zinv(A.shape[0], A.ctypes.data, out.ctypes.data)
    return out, info


def scba_mw(s, E: float, delta: float, coupling: np.ndarray, niter: int, tol: float,
            frm: Sequence[float], ref: int, eps_drop: float = 0.0, sigmas=None):
    """Returns dict(curr, iters, err, gn_diag, gr_diag)."""
    os_ = OracleStruct(s)
    cp = np.ascontiguousarray(coupling, dtype=np.float64)
    fr = np.ascontiguousarray(frm, dtype=np.float64)
    curr = np.zeros(s.ncont)
    err = np.zeros(1)
    gn = np.zeros(s.ndev, dtype=np.complex128)
    gr = np.zeros(s.ndev, dtype=np.complex128)
    if sigmas is None:
        it = lib().or_scba_meir_wingreen(os_.ptr, E, delta, cp.ctypes.data, niter, tol,
                                         fr.ctypes.data, ref, curr.ctypes.data, err.ctypes.data,
                                         gn.ctypes.data, gr.ctypes.data, SGFACC, eps_drop)
    else:
        sg = [_cz(x) for x in sigmas]
        arr = (C.c_void_p * len(sg))(*[x.ctypes.data for x in sg])
        it = lib().or_scba_mw_sigma(os_.ptr, E, C.cast(arr, C.c_void_p), cp.ctypes.data, niter,
                                    tol, fr.ctypes.data, ref, curr.ctypes.data, err.ctypes.data,
                                    gn.ctypes.data, gr.ctypes.data, eps_drop)
    return dict(curr=curr, iters=it, err=float(err[0]), gn_diag=gn, gr_diag=gr)


def gr_tridiag(s, E: float, sigmas: List[np.ndarray]):
    os_ = OracleStruct(s)
    sg = [_cz(x) for x in sigmas]
    arr = (C.c_void_p * len(sg))(*[x.ctypes.data for x in sg])
    b = s.bsz
    Grd = [np.zeros((n, n), dtype=np.complex128, order="F") for n in b]
    Gru = [np.zeros((b[i], b[i + 1]), dtype=np.complex128, order="F") for i in range(len(b) - 1)]
    Grl = [np.zeros((b[i + 1], b[i]), dtype=np.complex128, order="F") for i in range(len(b) - 1)]
    pd = (C.c_void_p * len(b))(*[x.ctypes.data for x in Grd])
    pu = (C.c_void_p * max(1, len(b) - 1))(*[x.ctypes.data for x in Gru])
    pl = (C.c_void_p * max(1, len(b) - 1))(*[x.ctypes.data for x in Grl])
    lib().or_gr_tridiag(os_.ptr, E, C.cast(arr, C.c_void_p), C.cast(pd, C.c_void_p),
                        C.cast(pu, C.c_void_p), C.cast(pl, C.c_void_p))
    return Grd, Gru, Grl


def fermi(E, Ef, kT):
    return lib().or_fermi(E, Ef, kT)


def integrate_el(T, mu1, mu2, kT1, kT2, emin, emax, estep):
    T = np.ascontiguousarray(T, dtype=np.float64)
    return lib().or_integrate_el(T.ctypes.data, len(T), mu1, mu2, kT1, kT2, emin, emax, estep)


def integrate_el_mw(integration, Iv, estep):
    Iv = np.ascontiguousarray(Iv, dtype=np.float64)
    return lib().or_integrate_el_mw(integration, Iv.ctypes.data, len(Iv), estep)


def npoints(emin, emax, estep):
    return lib().or_npoints(emin, emax, estep)


# ----------------------------------------------------------- density matrix --
def gauleg(x1: float, x2: float, n: int):
    x, w = np.zeros(n), np.zeros(n)
    lib().or_gauleg(x1, x2, x.ctypes.data, w.ctypes.data, n)
    return x, w


def contour_grid(Elow, muref, kbT, n_kt, np1, np2, npoles, g_spin):
    """contour_int_def (no inelastic terms): (z, w) complex arrays."""
    n = np1 + np2 + npoles
    z = np.zeros(max(n, 1), dtype=np.complex128)
    w = np.zeros(max(n, 1), dtype=np.complex128)
    lib().or_contour_grid(Elow, muref, kbT, n_kt, np1, np2, npoles, g_spin, z.ctypes.data, w.ctypes.data)
    return z[:n], w[:n]


def real_axis_grid(mumin, mumax, kbt_max, n_kt, npts, delta, g_spin):
    z = np.zeros(max(npts, 1), dtype=np.complex128)
    w = np.zeros(max(npts, 1), dtype=np.complex128)
    lib().or_real_axis_grid(mumin, mumax, kbt_max, n_kt, npts, delta, g_spin, z.ctypes.data, w.ctypes.data)
    return z[:npts], w[:npts]


def efa_contour_grid(particle, Eb, muref, kbT, n_kt, np1, np2, npoles, g_spin):
    """contour_int_n_def (particle 1) / contour_int_p_def (-1): (z, w)."""
    n = np1 + np2 + npoles
    z = np.zeros(max(n, 1), dtype=np.complex128)
    w = np.zeros(max(n, 1), dtype=np.complex128)
    k = lib().or_efa_contour_grid(particle, Eb, muref, kbT, n_kt, np1, np2, npoles, g_spin,
                                  z.ctypes.data, w.ctypes.data)
    return z[:k], w[:k]


def trapez(x1, x2, n):
    x, w = np.zeros(n), np.zeros(n)
    lib().or_trapez(x1, x2, x.ctypes.data, w.ctypes.data, n)
    return x, w


def density_efa(s, N, cstart, P, particle, p, outer=2):
    """compute_density_efa (libnegf.F90:1690-1791, no bulk contact density)
    from the parameters p (a dict of the lnparams fields used): the contour of
    contour_int_{n,p}_def, then the real-axis trapezoid of real_axis_int_{n,p}_def
    (integrations.F90:1346-1462) whose hole integrand is A - Gn
    (compute_Gp, :2405-2420) with A = i (Gr - Gr^H) of the G^r that comes
    with G^n (device at real E, contacts at E + i delta).  Returns (rho, pattern, q = Re diag rho)."""
    nc = len(p["mu"])
    mu = list(p["mu"])
    imin, imax = int(np.argmin(mu)), int(np.argmax(mu))
    ref = {0: imin + 1, 1: imax + 1}.get(p["min_or_max"], nc + 1)   # set_ref_cont
    muref = p["mu_n"][0] if particle == 1 else p["mu_p"][0]
    npts = p["np_n"] if particle == 1 else p["np_p"]
    contour = npts[0] + npts[1] + p["n_poles"] > 0
    if not contour:
        ref = nc + 1
    dore = p.get("dore", "D")
    rho = np.zeros((N, N), dtype=np.complex128)
    pat = np.eye(N, dtype=bool)
    if contour:
        kc = ref - 1 if ref <= nc else imax
        Eb = p["ec"] - p["deltaec"] if particle == 1 else p["ev"] + p["deltaev"]
        z, w = efa_contour_grid(particle, Eb, muref, p["kbt_dm"][kc], p["n_kt"], npts[0], npts[1],
                                p["n_poles"], p["g_spin"])
        if len(z):
            r, pt = density_matrix(s, N, cstart, P, contour=(z, w), outer=outer, dore=dore)
            rho, pat = rho + r, pat | pt
    if p["np_real"] > 0:
        Om = p["n_kt"] * max(p["kbt_dm"])
        if particle == 1:
            x, wt = trapez(p["ec"] - p["deltaec"], max(p["mu_n"]) + Om, p["np_real"])
        else:
            x, wt = trapez(min(p["mu_p"]) - Om, p["ev"] + p["deltaev"], p["np_real"])
        zr = x + 1j * p["delta"]
        wr = p["g_spin"] * wt / (2 * np.pi) * (x if dore == "E" else 1.0)
        frm_of = lambda E: [fermi(E, mu[c], p["kbt_dm"][c]) for c in range(nc)] + [0.0]
        g, pt = density_matrix(s, N, cstart, P, real=(zr, wr), frm_of=frm_of, ref=ref, outer=outer,
                               delta=p["delta"])
        if particle == 1:
            rho, pat = rho + g, pat | pt
        else:
            a, pa = density_matrix(s, N, cstart, P, contour=(zr, wr + 0j), outer=outer, real_device=True)
            rho, pat = rho + a - g, pat | pa | pt
    return rho, pat, np.real(np.diag(rho)).copy()


def density_quasi(s, particle, p, Ec, Ev, mun, mup):
    """quasiEq_int_n / quasiEq_int_p (integrations.F90:3002-3258), written
    loop for loop: per contour point the diagonal of G^r (device blocks; the
    contact rows carry no diagonal entry), weighted per orbital with its own
    quasi-Fermi level, then the poles of every orbital's level."""
    nz = len(Ec)
    kbT = max(p["kbt_dm"])
    Lam = 0.5 * kbT * np.pi if p["n_poles"] == 0 else 2.0 * p["n_poles"] * kbT * np.pi
    Om = p["n_kt"] * kbT
    kw = p.get("kwght", 1.0)
    g = p["g_spin"]
    if particle == 1:
        minE = np.asarray(Ec) - p["deltaec"]
        maxE = np.asarray(Ec) + Om
        segs = [(minE.min() + 1j * Lam, maxE.max() + 1j * Lam, p["np_n"][1]),
                (minE.min() + 0j, minE.min() + 1j * Lam, p["np_n"][0])]
        lev, sgn = np.asarray(mun), 1.0
    else:
        minE = np.asarray(Ev) - Om
        maxE = np.asarray(Ev) + p["deltaev"]
        segs = [(maxE.max() + 1j * Lam, minE.min() + 1j * Lam, p["np_p"][1]),
                (maxE.max() + 0j, maxE.max() + 1j * Lam, p["np_p"][0])]
        lev, sgn = np.asarray(mup), -1.0
    bs = s.bsz
    off = np.concatenate([[0], np.cumsum(bs)])

    def diag(z):
        G = gr_point(s, complex(z))["Gd"]
        d = np.zeros(nz, dtype=np.complex128)
        for k in range(min(nz, off[-1])):
            b = int(np.searchsorted(off, k, side="right")) - 1
            d[k] = G[b][k - off[b], k - off[b]]
        return d

    def ferm(z, mu):
        return 1.0 / (np.exp((z - mu) / kbT) + 1.0)

    temp = np.zeros(nz, dtype=np.complex128)
    for a, b, n in segs:
        x, w = gauleg(0.0, 1.0, n)
        zd = b - a
        for i in range(n):
            Ez = a + x[i] * zd
            d = diag(Ez)
            ff = ferm(Ez, lev) if particle == 1 else ferm(-Ez, -lev)
            temp += d * (sgn * g * kw * w[i] * zd * ff / (2 * np.pi))
    for k in range(nz):
        for i in range(1, p["n_poles"] + 1):
            Ez = lev[k] + 1j * kbT * np.pi * (2 * i - 1)
            temp[k] += diag(Ez)[k] * (-sgn * 1j * kbT * g)
    return np.real(1j * (temp - np.conj(temp)))


def _blocks(shapes):
    arrs = [np.zeros(sh, dtype=np.complex128, order="F") for sh in shapes]
    ptr = (C.c_void_p * max(len(arrs), 1))(*[a.ctypes.data for a in arrs])
    return arrs, C.cast(ptr, C.c_void_p), ptr


def _point_shapes(s):
    b = s.bsz
    nd = [b[c.cblk - 1] for c in s.contacts]
    nc = [c.nc + c.ns for c in s.contacts]  # outer blocks span surface + PL1
    return ([(n, n) for n in b], [(b[i], b[i + 1]) for i in range(len(b) - 1)],
            [(b[i + 1], b[i]) for i in range(len(b) - 1)], [(nd[k], nc[k]) for k in range(len(nc))],
            [(nc[k], nd[k]) for k in range(len(nc))])


def gr_point(s, z: complex, real_device: bool = False):
    """G^r blocks at complex z with the outer blocks (unmasked); with
    real_device the device is at Re z and only the contacts at z."""
    os_ = OracleStruct(s)
    shd, shu, shl, shlc, shcl = _point_shapes(s)
    Gd, pd, kd = _blocks(shd)
    Gu, pu, ku = _blocks(shu)
    Gl, pl, kl = _blocks(shl)
    Glc, plc, klc = _blocks(shlc)
    Gcl, pcl, kcl = _blocks(shcl)
    ncyc = np.zeros(max(s.ncont, 1), dtype=np.int32)
    fn = lib().or_gr_point_real if real_device else lib().or_gr_point
    fn(os_.ptr, z.real, z.imag, pd, pu, pl, plc, pcl, ncyc.ctypes.data, SGFACC)
    return dict(Gd=Gd, Gu=Gu, Gl=Gl, Glc=Glc, Gcl=Gcl, ncyc=ncyc[:s.ncont])


def gn_point(s, E: float, delta: float, frm, ref: int, eps: float = 0.0):
    """G^n blocks on the real axis with the masked outer blocks."""
    os_ = OracleStruct(s)
    shd, shu, shl, shlc, shcl = _point_shapes(s)
    Gd, pd, kd = _blocks(shd)
    Gu, pu, ku = _blocks(shu)
    Gl, pl, kl = _blocks(shl)
    Glc, plc, klc = _blocks(shlc)
    Gcl, pcl, kcl = _blocks(shcl)
    fr = np.ascontiguousarray(frm, dtype=np.float64)
    ncyc = np.zeros(max(s.ncont, 1), dtype=np.int32)
    lib().or_gn_point(os_.ptr, E, delta, fr.ctypes.data, ref, pd, pu, pl, plc, pcl, ncyc.ctypes.data,
                      SGFACC, eps)
    return dict(Gd=Gd, Gu=Gu, Gl=Gl, Glc=Glc, Gcl=Gcl, ncyc=ncyc[:s.ncont])


def _t_pattern(s, k):
    c = s.contacts[k]
    p = c.HMC != 0
    if c.SMC is not None:
        p = p | (c.SMC != 0)
    return p


def _add_point(s, N, cstart, P, T, pat, blk, w, lower, masked_outer, tl_pattern=None):
    """Adds w * blocks onto the pattern: the device part through blk2csr
    (iterative.F90:1375-1450) on the S pattern P (rows/columns of the device
    only), the outer blocks at (block cb rows, contact surface + PL1 columns)
    and, with lower, at the transposed position (calculate_Gr_outer /
    calculate_Gn_outer)."""
    off = s.block_offsets()
    nd = s.ndev
    rr, cc = np.nonzero(P[:nd, :nd])
    bidx = np.searchsorted(off, np.arange(nd), side="right") - 1
    for r, c in zip(rr, cc):
        x, y = bidx[r], bidx[c]
        if y == x:
            v = blk["Gd"][x][r - off[x], c - off[y]]
        elif y == x + 1:
            v = blk["Gu"][x][r - off[x], c - off[y]]
        elif y == x - 1:
            v = blk["Gl"][y][r - off[x], c - off[y]]
        else:
            raise ValueError("S pattern outside the block tridiagonal (blk2csr)")
        T[r, c] += w * v
    pat |= P
    for k, con in enumerate(s.contacts):
        cb = con.cblk - 1
        r0, c0, nc, ndk = off[cb], cstart[k], con.nc + con.ns, s.bsz[cb]
        tp = _t_pattern(s, k)
        up = blk["Glc"][k]
        m = tp & (up != 0)
        sub = T[r0:r0 + ndk, c0:c0 + nc]
        sub[m] += w * up[m]
        pat[r0:r0 + ndk, c0:c0 + nc] |= m
        if lower:
            lo = blk["Gcl"][k]
            ml = (tp.T if not masked_outer else m.T) & (lo != 0)
            sub = T[c0:c0 + nc, r0:r0 + ndk]
            sub[ml] += w * lo[ml]
            pat[c0:c0 + nc, r0:r0 + ndk] |= ml


def density_matrix(s, N, cstart, P, contour=None, real=None, frm_of=None, ref=None, outer=2,
                   dore="D", delta=1e-4, eps=0.0, real_device=False):
    """Dense N x N density matrix and its pattern as compute_density_dft builds
    it (contour_int, integrations.F90:844-897; real_axis_int, :1093-1176):
    T = sum_i w_i Gr(z_i) on the pattern, rho = i (T - T^H), then
    rho += sum_j w_j Gn(E_j).  P: N x N boolean S pattern; cstart[k]: 0-based
    first column of contact k (its surface block, else its first PL);
    frm_of(E) -> frm[ncont+1]."""
    rho = np.zeros((N, N), dtype=np.complex128)
    pat = np.zeros((N, N), dtype=bool)
    pat[np.arange(N), np.arange(N)] = True
    if contour is not None:
        T = np.zeros((N, N), dtype=np.complex128)
        pt = np.eye(N, dtype=bool)
        for z, w in zip(*contour):
            blk = gr_point(s, complex(z), real_device)
            ww = w * z if dore == "E" else w
            _add_point(s, N, cstart, P, T, pt, blk, ww, outer == 2, False)
        rho = 1j * (T - T.conj().T)
        pat = pt | pt.T
    if real is not None:
        T = np.zeros((N, N), dtype=np.complex128)
        pt = np.eye(N, dtype=bool)
        for z, w in zip(*real):
            E = float(np.real(z))
            blk = gn_point(s, E, float(np.imag(z)), frm_of(E), ref, eps)
            ww = w * E if dore == "E" else w
            _add_point(s, N, cstart, P, T, pt, blk, ww, outer == 2, True)
        rho = rho + T
        pat = pat | pt
    return rho, pat


# ---- inelastic electron-phonon SCBA (elphinel.F90, self_energy.F90) ----------
_ZERO_LATT = np.zeros(9)


def inel_kmat_po(ndz: int, dz: float, kfrac: np.ndarray, kw: np.ndarray, coupling: float,
                 wq: float, area: float, eps0: float, eps_inf: float, q0: float,
                 latt: Optional[np.ndarray] = None) -> np.ndarray:
    """prepare_POKmat: K[iz, iK, iQ] (returned as an (ndz, nk, nk) Fortran-order array)."""
    kf = np.ascontiguousarray(np.asarray(kfrac, dtype=np.float64).reshape(-1, 3))
    nk = kf.shape[0]
    w = np.ascontiguousarray(kw, dtype=np.float64)
    la = np.ascontiguousarray(_ZERO_LATT if latt is None else np.asarray(latt, float).T.ravel())
    K = np.zeros(ndz * nk * nk)
    L = lib()
    L.or_inel_kmat_po.argtypes = [C.c_int, C.c_double, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p] + \
        [C.c_double] * 6 + [C.c_void_p]
    L.or_inel_kmat_po(ndz, dz, nk, kf.ctypes.data, w.ctypes.data, la.ctypes.data, coupling, wq, area,
                      eps0, eps_inf, q0, K.ctypes.data)
    return K.reshape((ndz, nk, nk), order="F")


def inel_kmat_nonpo(ndz: int, dz: float, kfrac: np.ndarray, kw: np.ndarray, coupling: float,
                    D0: float, latt: Optional[np.ndarray] = None) -> np.ndarray:
    """prepare_NonPOKmat: K[iz, iK, iQ]."""
    kf = np.ascontiguousarray(np.asarray(kfrac, dtype=np.float64).reshape(-1, 3))
    nk = kf.shape[0]
    w = np.ascontiguousarray(kw, dtype=np.float64)
    la = np.ascontiguousarray(_ZERO_LATT if latt is None else np.asarray(latt, float).T.ravel())
    K = np.zeros(ndz * nk * nk)
    L = lib()
    L.or_inel_kmat_nonpo.argtypes = [C.c_int, C.c_double, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                                     C.c_double, C.c_double, C.c_void_p]
    L.or_inel_kmat_nonpo(ndz, dz, nk, kf.ctypes.data, w.ctypes.data, la.ctypes.data, coupling, D0,
                         K.ctypes.data)
    return K.reshape((ndz, nk, nk), order="F")


def _blk_sizes(bsz):
    nd = sum(b * b for b in bsz)
    no = sum(bsz[i] * bsz[i + 1] for i in range(len(bsz) - 1))
    return nd, no


def inelastic_mw(structs, kw, E: np.ndarray, delta: float, frm: np.ndarray, ref: int,
                 K: np.ndarray, iz: np.ndarray, shift: int, Nq: float, tridiag: bool,
                 niter: int, tol: float, coupling: Optional[np.ndarray] = None, niter_ela: int = 0,
                 tol_ela: float = 1e-7, integration: int = 0, g_spin: float = 2.0,
                 stores: bool = False, nthreads: int = 0, eps_drop: float = 0.0) -> dict:
    """meir_wingreen with an inelastic model.  structs[nk]: the device at each
    k-point; E: uniform grid; frm[ne, ncont+1].  Returns dict(curr_mat[ne, ncont],
    currents[ncont], passes, err, iters_ela[nk, ne]) and, with stores=True, the
    per-point block stores (flat, point p = iK*ne + iE)."""
    structs = list(structs)
    nk, ne = len(structs), len(E)
    keep = [OracleStruct(s) for s in structs]
    arr = (C.c_void_p * nk)(*[C.cast(C.pointer(k.c), C.c_void_p) for k in keep])
    s0 = structs[0]
    ncont = s0.ncont
    Ev = np.ascontiguousarray(E, dtype=np.float64)
    w = np.ascontiguousarray(kw, dtype=np.float64)
    fr = np.ascontiguousarray(np.asarray(frm, dtype=np.float64).reshape(ne, ncont + 1))
    Kc = np.ascontiguousarray(np.asarray(K, dtype=np.float64).ravel(order="F"))
    ndz = np.asarray(K).shape[0]
    izc = np.ascontiguousarray(iz, dtype=np.int32)
    cp = None if coupling is None else np.ascontiguousarray(coupling, dtype=np.float64)
    cm = np.zeros(ne * ncont)
    cur = np.zeros(ncont)
    err = np.zeros(1)
    its = np.zeros(ne * nk, dtype=np.int32)
    nd, no = _blk_sizes(list(s0.bsz))
    st = None
    sp = None
    if stores:
        st = [np.zeros(ne * nk * (nd if t in (0, 3, 6, 9) else max(no, 1)), dtype=np.complex128)
              for t in range(11)]
        sp = (C.c_void_p * 11)(*[x.ctypes.data for x in st])
    L = lib()
    L.or_inel_meir_wingreen.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_double,
                                        C.c_void_p, C.c_int, C.c_double, C.c_void_p, C.c_int,
                                        C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_double, C.c_int,
                                        C.c_int, C.c_double, C.c_int, C.c_double, C.c_double,
                                        C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                        C.c_double, C.c_double, C.c_int]
    estep = float(E[1] - E[0]) if ne > 1 else 0.0
    p = L.or_inel_meir_wingreen(nk, C.cast(arr, C.c_void_p), w.ctypes.data, ne, Ev.ctypes.data, delta,
                                None if cp is None else cp.ctypes.data, niter_ela, tol_ela,
                                fr.ctypes.data, ref, Kc.ctypes.data, ndz, izc.ctypes.data, shift, Nq,
                                1 if tridiag else 0, niter, tol, integration, estep, g_spin,
                                cm.ctypes.data, cur.ctypes.data, err.ctypes.data, its.ctypes.data,
                                None if sp is None else C.cast(sp, C.c_void_p), SGFACC, eps_drop, nthreads)
    if p < 0:
        raise RuntimeError(f"or_inel_meir_wingreen failed ({p})")
    out = dict(curr_mat=cm.reshape((ne, ncont), order="F"), currents=cur, passes=p,
               err=float(err[0]), iters_ela=its.reshape(nk, ne))
    if stores:
        out["stores"] = dict(zip(["Grd", "Gru", "Grl", "Gnd", "Gnu", "Gnl", "Srd", "Sru", "Srl", "Snd", "Snu"], st))
    return out


def inel_sigma(bsz, ne: int, nk: int, K: np.ndarray, iz: np.ndarray, shift: int, Nq: float,
               tridiag: bool, G: Sequence[np.ndarray]) -> List[np.ndarray]:
    """compute_Sigma_n + compute_Sigma_r from the per-point G stores
    G = [Grd, Gru, Grl, Gnd, Gnu, Gnl]; returns [Srd, Sru, Srl, Snd, Snu]."""
    nd, no = _blk_sizes(list(bsz))
    out = [np.zeros(ne * nk * (nd if t in (0, 3) else max(no, 1)), dtype=np.complex128) for t in range(5)]
    b = np.array(bsz, dtype=np.int32)
    Kc = np.ascontiguousarray(np.asarray(K, dtype=np.float64).ravel(order="F"))
    izc = np.ascontiguousarray(iz, dtype=np.int32)
    g = [np.ascontiguousarray(x, dtype=np.complex128) for x in G]
    L = lib()
    L.or_inel_sigma.argtypes = [C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_void_p,
                                C.c_int, C.c_double, C.c_int] + [C.c_void_p] * 11
    rc = L.or_inel_sigma(len(b), b.ctypes.data, ne, nk, Kc.ctypes.data, np.asarray(K).shape[0],
                         izc.ctypes.data, shift, Nq, 1 if tridiag else 0, *[x.ctypes.data for x in g],
                         *[x.ctypes.data for x in out])
    if rc:
        raise RuntimeError(f"or_inel_sigma failed ({rc})")
    return out
