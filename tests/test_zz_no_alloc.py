"""hipnegf_Zinverse_noalloc (the allocation-free block inverse) against
hipnegf_Zinverse through the C ABI: the same kernels, so the same inverse, on
two calls in a row (a blocked-path block, then a smaller one), plus the size
limit of its scratch array."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rand(n, seed):
    rng = np.random.default_rng(seed)
    a = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    return np.asfortranarray(a + 0.5 * n ** 0.5 * np.eye(n))


def test_noalloc_inverse_matches_zinverse():
    from libnegf_amd import hip
    L = hip.load()
    h = C.c_void_p()
    assert L.hipnegf_blasInit(C.byref(h)) == 0, L.hipnegf_last_error()
    bufs = []

    def dmat(nbytes):
        p = C.c_void_p()
        assert L.hipnegf_createMat(C.byref(p), nbytes) == 0, L.hipnegf_last_error()
        bufs.append(p)
        return p

    try:
        nmax = 400
        dA, dX, dY = dmat(nmax * nmax * 16), dmat(nmax * nmax * 16), dmat(nmax * nmax * 16)
        for n, seed in ((400, 1), (72, 2)):  # second call smaller (and on the LDS path)
            A = _rand(n, seed)
            assert L.hipnegf_copyMatH2D(A.ctypes.data, dA, A.nbytes) == 0
            rc1 = L.hipnegf_Zinverse(h, None, dA, dX, n)
            rc2 = L.hipnegf_Zinverse_noalloc(h, dA, dY, n)
            assert rc1 == 0 and rc2 == 0, L.hipnegf_last_error()
            X = np.zeros((n, n), np.complex128, order="F")
            Y = np.zeros((n, n), np.complex128, order="F")
            A2 = np.zeros_like(Y)
            assert L.hipnegf_copyMatD2H(X.ctypes.data, dX, X.nbytes) == 0
            assert L.hipnegf_copyMatD2H(Y.ctypes.data, dY, Y.nbytes) == 0
            assert L.hipnegf_copyMatD2H(A2.ctypes.data, dA, A2.nbytes) == 0
            assert np.array_equal(A2, A)                         # A unchanged
            assert np.abs(Y - X).max() <= 1e-13 * np.abs(X).max(), n   # same kernels
            ref = np.linalg.inv(A)
            assert np.abs(Y - ref).max() <= 1e-11 * np.abs(ref).max(), n
        # the scratch array holds one block of order <= 512
        assert L.hipnegf_Zinverse_noalloc(h, dA, dY, 513) == -1
        assert b"512" in L.hipnegf_last_error()
    finally:
        for p in bufs:
            L.hipnegf_deleteMat(C.byref(p))
        L.hipnegf_blasFinalize(h)
