"""Python front end of the libNEGF C API on the MI355X engine.

Mirrors the reference's ctypes class ``NEGF`` (tests/cpython/libnegf.py):
same method names, argument meaning and call order (``set_h`` from a scipy
CSR matrix, ``set_s_id``, ``init_structure``, ``params`` edited in place then
``set_params``, ``solve_landauer``, ``get_energies`` / ``get_currents`` /
``get_transmission``), bound to ``libhipnegf.so`` (include/hipnegf_api.h)
instead of the Fortran library.  The parameter block follows the current C
header layout (src/api/lnParams.h:15-55).

Differences in error behaviour: the reference stops the process on an error;
here the call records it and this class raises ``NegfError`` with the
library's message.  There is no CPU fallback: the HIP library must be built.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np

from . import hip

MAXCONT = 10
MAXLAYERS = 10000


class NegfError(RuntimeError):
    pass


class LNParams(C.Structure):
    """struct lnparams (src/api/lnParams.h:15-55)."""
    _fields_ = [
        ("verbose", C.c_int),
        ("readold_dm_sgfs", C.c_int),
        ("readold_t_sgfs", C.c_int),
        ("sgfcache", C.c_int),
        ("spin", C.c_int),
        ("kpoint", C.c_int),
        ("g_spin", C.c_double),
        ("delta", C.c_double),
        ("deltaModel", C.c_int),
        ("wmax", C.c_double),
        ("dos_delta", C.c_double),
        ("eneconv", C.c_double),
        ("kwght", C.c_double),
        ("ec", C.c_double),
        ("ev", C.c_double),
        ("deltaec", C.c_double),
        ("deltaev", C.c_double),
        ("emin", C.c_double),
        ("emax", C.c_double),
        ("estep", C.c_double),
        ("estep_coarse", C.c_double),
        ("mu_n", C.c_double * MAXCONT),
        ("mu_p", C.c_double * MAXCONT),
        ("mu", C.c_double * MAXCONT),
        ("contact_dos", C.c_double * MAXCONT),
        ("fictcont", C.c_bool * MAXCONT),
        ("kbt_dm", C.c_double * MAXCONT),
        ("kbt_t", C.c_double * MAXCONT),
        ("scba_inelastic_tol", C.c_double),
        ("scba_elastic_tol", C.c_double),
        ("np_n", C.c_int * 2),
        ("np_p", C.c_int * 2),
        ("np_real", C.c_int),
        ("n_kt", C.c_int),
        ("n_poles", C.c_int),
        ("ni", C.c_int * MAXLAYERS),
        ("nf", C.c_int * MAXLAYERS),
        ("dore", C.c_char * 1),
        ("min_or_max", C.c_int),
        ("is_s_is", C.c_bool),
    ]


def _i32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.int32)


class NEGF:
    """One libNEGF session (negf_init_session + negf_init)."""

    def __init__(self, path: Optional[str] = None):
        self._lib = hip.load() if path is None else C.CDLL(path)
        L = self._lib
        hs = C.c_int()
        L.negf_gethandlersize(C.byref(hs))
        self._handler = (C.c_int * hs.value)()
        L.negf_get_last_error.restype = C.c_char_p
        L.negf_init_session(self._handler)
        L.negf_init(self._handler)
        self._check()
        self.params = LNParams()
        L.negf_get_params(self._handler, C.byref(self.params))
        self._closed = False

    # ---- plumbing --------------------------------------------------------
    def _check(self):
        if self._lib.negf_get_status(self._handler) != 0:
            raise NegfError(self._lib.negf_get_last_error(self._handler).decode())

    def _call(self, name, *args):
        getattr(self._lib, name)(self._handler, *args)
        self._check()

    def close(self):
        if not getattr(self, "_closed", True):
            self._lib.negf_destruct_libnegf(self._handler)
            self._lib.negf_destruct_session(self._handler)
            self._closed = True

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- parameters ------------------------------------------------------
    def get_params(self):
        self._call("negf_get_params", C.byref(self.params))

    def set_params(self):
        self._call("negf_set_params", C.byref(self.params))

    def set_verbosity(self, level: int):
        self._call("negf_set_verbosity", C.c_int(level))

    # ---- system ----------------------------------------------------------
    def create_hs(self, nhk: int):
        self._call("negf_create_hs", C.c_int(nhk))

    @staticmethod
    def _csr_parts(mat):
        mat = mat.tocsr()
        mat.sort_indices()
        re = np.ascontiguousarray(np.real(mat.data), dtype=np.float64)
        im = np.ascontiguousarray(np.imag(mat.data), dtype=np.float64)
        return mat.shape[0], re, im, _i32(mat.indices + 1), _i32(mat.indptr + 1)

    def set_h(self, mat, ik: Optional[int] = None):
        """H from a scipy sparse matrix (PL-ordered rows, as the reference requires)."""
        n, re, im, ja, ia = self._csr_parts(mat)
        if ik is None:
            self._call("negf_set_h_cp", C.c_int(n), hip._ptr(re), hip._ptr(im), hip._ptr(ja), hip._ptr(ia))
        else:
            z = np.ascontiguousarray(re + 1j * im)
            self._call("negf_set_h", C.c_int(n), hip._ptr(z), hip._ptr(ja), hip._ptr(ia), C.c_int(ik))

    def set_s(self, mat, ik: Optional[int] = None):
        n, re, im, ja, ia = self._csr_parts(mat)
        if ik is None:
            self._call("negf_set_s_cp", C.c_int(n), hip._ptr(re), hip._ptr(im), hip._ptr(ja), hip._ptr(ia))
        else:
            z = np.ascontiguousarray(re + 1j * im)
            self._call("negf_set_s", C.c_int(n), hip._ptr(z), hip._ptr(ja), hip._ptr(ia), C.c_int(ik))

    def set_s_id(self, nrow: int, ik: int = 1):
        self._call("negf_set_s_id", C.c_int(nrow), C.c_int(ik))

    def set_kpoint(self, ik: int):
        self._call("negf_set_kpoint", C.c_int(ik))

    def init_structure(self, ncont, contend, surfend, npl, plend, cblk, surfstart=None):
        """Same argument order as the reference's Python class; ``surfstart``
        defaults to surfend + 1 (no surface block)."""
        contend, surfend, plend, cblk = _i32(contend), _i32(surfend), _i32(plend), _i32(cblk)
        surfstart = surfend + 1 if surfstart is None else _i32(surfstart)
        self._call("negf_init_contacts", C.c_int(ncont))
        self.ncont = int(ncont)
        self._call("negf_init_structure", C.c_int(ncont), hip._ptr(surfstart), hip._ptr(surfend),
                   hip._ptr(contend), C.c_int(npl), hip._ptr(plend), hip._ptr(cblk))
        self.get_params()

    def contact_blocks(self, ncont, surfstart, surfend, contend, npl, plend):
        """Device PL each contact couples to, found from H (``negf_contact_blocks``)."""
        surfstart, surfend, contend, plend = _i32(surfstart), _i32(surfend), _i32(contend), _i32(plend)
        cblk = np.zeros(ncont, dtype=np.int32)
        self._call("negf_contact_blocks", C.c_int(ncont), hip._ptr(surfstart), hip._ptr(surfend),
                   hip._ptr(contend), C.c_int(npl), hip._ptr(plend), hip._ptr(cblk))
        return cblk

    def get_pls(self):
        """(plend, cblk) of the structure (``negf_get_pls``)."""
        npl, nc = C.c_int(), C.c_int()
        d = np.zeros(1, dtype=np.int32)
        self._call("negf_get_pls", C.byref(npl), C.byref(nc), hip._ptr(d), hip._ptr(d), C.c_int(0))
        plend = np.zeros(max(npl.value, 1), dtype=np.int32)
        cblk = np.zeros(MAXCONT, dtype=np.int32)
        self._call("negf_get_pls", C.byref(npl), C.byref(nc), hip._ptr(plend), hip._ptr(cblk), C.c_int(1))
        return plend[:npl.value], cblk[:nc.value]

    def set_reference(self, minmax: int):
        """Reference contact of the density paths: 0 lowest mu, 1 highest."""
        v = C.c_int(minmax)
        self._call("negf_set_reference", C.byref(v))
        self.params.min_or_max = minmax

    def get_version(self):
        self._call("negf_get_version")

    def print_mat(self):
        """H.dat and S.dat into the scratch path (printcsr format)."""
        self._call("negf_print_mat")

    def write_partition(self):
        self._call("negf_write_partition")

    def print_tnegf(self):
        self.set_params()
        self._call("negf_print_tnegf")

    def mem_stats(self):
        self._call("negf_mem_stats")

    def set_elph_dephasing(self, coupling, niter: int):
        cp = np.ascontiguousarray(coupling, dtype=np.float64)
        self._call("negf_set_elph_dephasing", hip._ptr(cp), C.c_int(cp.size), C.c_int(niter))

    def set_scba_tolerances(self, elastic_tol: float, inelastic_tol: float):
        self._call("negf_set_scba_tolerances", C.c_double(elastic_tol), C.c_double(inelastic_tol))

    def set_precision(self, precision: int):
        self._call("negf_set_precision", C.c_int(precision))

    def set_batch(self, max_batch: int):
        self._call("negf_set_batch", C.c_int(max_batch))

    # ---- solves ----------------------------------------------------------
    def solve_landauer(self):
        """compute_current: Landauer transmission (or the dephasing effective
        transmission when a dephasing model is set) and the integrated currents."""
        self.set_params()
        self._call("negf_solve_landauer")

    def calculate_dephasing_transmission(self):
        self.set_params()
        self._call("negf_calculate_dephasing_transmission")

    def read_hs(self, re_fname: str, im_fname: str, target: int):
        """H (target 0) or S (target 1) from the PETSc ASCII real/imaginary files."""
        self._call("negf_read_hs", C.c_char_p(re_fname.encode()), C.c_char_p(im_fname.encode()),
                   C.c_int(target))

    def set_output(self, path: str):
        self._call("negf_set_output", C.c_char_p(path.encode()))

    def set_scratch(self, path: str):
        """Folder of the surface-function files (``negf_set_scratch``,
        libnegf_api.f90:314-334); used when ``params.sgfcache == 0``."""
        self._call("negf_set_scratch", C.c_char_p(path.encode()))

    def write_tun_and_dos(self):
        """tunneling_*.dat and localDOS_*.dat in the output path."""
        self._call("negf_write_tunneling_and_dos")

    def solve_density(self):
        """compute_density_dft: the density matrix by contour (and, with
        params.np_real > 0, real-axis) integration; read it with get_dm()."""
        self.set_params()
        self._call("negf_solve_density_dft")

    def density_efa(self, ndofs: int, particle: int, contacts: bool = False):
        """compute_density_efa: electron (particle 1) or hole (-1) density per
        orbital, Re diag(rho) of the first ndofs orbitals."""
        self.set_params()
        q = np.zeros(ndofs)
        self._call("negf_density_efa", C.byref(C.c_int(ndofs)), hip._ptr(q), C.byref(C.c_int(particle)),
                   C.byref(C.c_bool(contacts)))
        return q

    def density_quasi_eq(self, particle: int, Ec, Ev, mu_n, mu_p):
        """compute_density_quasiEq: electron (1) or hole (-1) density of each
        orbital for band-edge profiles Ec/Ev and quasi-Fermi levels mu_n/mu_p."""
        self.set_params()
        Ec, Ev, mu_n, mu_p = (np.ascontiguousarray(a, dtype=np.float64) for a in (Ec, Ev, mu_n, mu_p))
        q = np.zeros(Ec.size)
        self._call("negf_density_quasi_equilibrium", C.byref(C.c_int(Ec.size)), hip._ptr(q),
                   C.byref(C.c_int(particle)), hip._ptr(Ec), hip._ptr(Ev), hip._ptr(mu_n), hip._ptr(mu_p))
        return q

    def set_outer(self, outer: int):
        """0: device blocks only, 1: + device/contact, 2: + contact/device (default)."""
        self._call("negf_set_outer", C.byref(C.c_int(outer)))

    def get_dm(self):
        """The density matrix as a scipy CSR matrix (complex)."""
        import scipy.sparse as sps
        nnz, nrow = C.c_int(), C.c_int()
        d = np.zeros(1, dtype=np.int32)
        dd = np.zeros(1)
        self._call("negf_get_dm", C.byref(nnz), C.byref(nrow), hip._ptr(d), hip._ptr(d), hip._ptr(dd),
                   hip._ptr(dd), C.c_int(0))
        rowpnt = np.zeros(nrow.value + 1, dtype=np.int32)
        colind = np.zeros(max(nnz.value, 1), dtype=np.int32)
        re = np.zeros(max(nnz.value, 1))
        im = np.zeros(max(nnz.value, 1))
        self._call("negf_get_dm", C.byref(nnz), C.byref(nrow), hip._ptr(rowpnt), hip._ptr(colind),
                   hip._ptr(re), hip._ptr(im), C.c_int(1))
        n = nnz.value
        return sps.csr_matrix(((re + 1j * im)[:n], colind[:n] - 1, rowpnt - 1),
                              shape=(nrow.value, nrow.value), dtype=np.complex128)

    def current(self, units_h: str = "unknown", units_j: str = "unknown") -> float:
        self.set_params()
        out = C.c_double()
        self._call("negf_current", C.byref(out), C.c_char_p(units_h.encode()), C.c_char_p(units_j.encode()))
        return out.value

    # ---- results ---------------------------------------------------------
    def get_energies(self):
        n = C.c_int()
        dummy = np.zeros(1)
        self._call("negf_get_energies", C.byref(n), hip._ptr(dummy), hip._ptr(dummy), C.c_int(0))
        re = np.zeros(n.value)
        im = np.zeros(n.value)
        self._call("negf_get_energies", C.byref(n), hip._ptr(re), hip._ptr(im), C.c_int(1))
        return re, im

    def get_currents(self):
        n = C.c_int()
        dummy = np.zeros(1)
        self._call("negf_get_currents", C.byref(n), hip._ptr(dummy), C.c_int(0))
        out = np.zeros(n.value)
        self._call("negf_get_currents", C.byref(n), hip._ptr(out), C.c_int(1))
        return out

    def _associated(self, name):
        shape = (C.c_int * 2)()
        p = C.POINTER(C.c_double)()
        self._call(name, shape, C.byref(p))
        if not p or shape[0] == 0 or shape[1] == 0:
            return np.zeros((0, 0))
        # column-major (npoints, ncol) -> rows are columns
        return np.ctypeslib.as_array(p, shape=(shape[1], shape[0])).copy()

    def get_transmission(self):
        """(npairs, npoints): one row per lead pair, as the reference's class."""
        return self._associated("negf_associate_transmission")

    def compute_ldos(self):
        """compute_ldos: ldos_mat from -g_spin Im diag(G^r S) / pi at E + i dos_delta
        on the transmission grid (libnegf.F90:1893-1901)."""
        self.set_params()
        self._call("negf_compute_ldos")

    def get_ldos(self):
        """(nldos, npoints): local DOS of each projection, as the reference's class."""
        return self._associated("negf_associate_ldos")

    def set_ldos_intervals(self, istart, iend):
        """LDOS on orbital intervals istart[i]..iend[i] (1-based, inclusive)."""
        a, b = _i32(istart), _i32(iend)
        self._call("negf_init_ldos", C.c_int(a.size))
        self._call("negf_set_ldos_intervals", C.c_int(a.size), hip._ptr(a), hip._ptr(b))

    def init_ldos(self, nldos: int):
        self._call("negf_init_ldos", C.c_int(nldos))

    def set_ldos_indexes(self, ildos: int, idx):
        """Projection ildos (1-based) sums the given 1-based orbitals."""
        a = _i32(idx)
        self._call("negf_set_ldos_indexes", C.c_int(ildos), C.c_int(a.size), hip._ptr(a))

    def set_tun_indexes(self, idx=None):
        """tun_proj: only these 1-based orbitals of the traced block enter T (None: all)."""
        if idx is None:
            self._call("negf_set_tun_indexes", C.c_int(-1), None)
        else:
            a = _i32(idx)
            self._call("negf_set_tun_indexes", C.c_int(a.size), hip._ptr(a))

    def get_energy_current(self):
        """(ncont, npoints) Meir-Wingreen energy-resolved currents."""
        return self._associated("negf_associate_energy_current")

    def get_item_status(self):
        n = C.c_int()
        d = np.zeros(1, dtype=np.int32)
        self._call("negf_get_item_status", C.byref(n), hip._ptr(d), None, C.c_int(0))
        st = np.zeros(n.value, dtype=np.int32)
        self._call("negf_get_item_status", C.byref(n), hip._ptr(st), None, C.c_int(1))
        return st

    def get_item_ncyc(self):
        """Decimation iteration counts of the last solve, npoints x ncont
        entries (0 for a surface function loaded from a cache)."""
        n = C.c_int()
        d = np.zeros(1, dtype=np.int32)
        self._call("negf_get_item_status", C.byref(n), hip._ptr(d), None, C.c_int(0))
        st = np.zeros(n.value, dtype=np.int32)
        nc = np.zeros(n.value * max(getattr(self, "ncont", 0), 1), dtype=np.int32)
        self._call("negf_get_item_status", C.byref(n), hip._ptr(st), hip._ptr(nc), C.c_int(1))
        return nc
