"""GPU unit tests of the batched MFMA complex GEMM and the batched block inverse
(C ABI: hipnegf_{z,c}gemm_batched, hipnegf_{z,c}inverse_batched).

References: numpy (LAPACK/BLAS) in complex128.  GEMM tolerance 1e-13 of the
operand-norm bound in FP64 (4e-6 in FP32); inverses checked through the
residual ||A X - I|| / (||A|| ||X||) <= 1e-13 (FP64), 1e-5 (FP32).
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from libnegf_amd import hip
    eng = hip.Engine(0)
    return torch, hip, eng


def _rand(rng, shape):
    return rng.normal(size=shape) + 1j * rng.normal(size=shape)


OPS = {"N": 0, "T": 1, "C": 2}


def _op(a, o):
    return a if o == "N" else (a.T if o == "T" else a.conj().T)


@pytest.mark.parametrize("m,n,k", [(16, 16, 4), (64, 64, 64), (70, 33, 45), (1, 1, 1),
                                   (129, 65, 200), (256, 256, 256)])
@pytest.mark.parametrize("ops", ["NN", "CN", "NC", "TN"])
def test_zgemm_batched(gpu, m, n, k, ops):
    torch, hip, eng = gpu
    rng = np.random.default_rng(m * 1000 + n * 10 + k)
    batch = 3
    oa, ob = ops
    A = [_rand(rng, (m, k) if oa == "N" else (k, m)) for _ in range(batch)]
    B = [_rand(rng, (k, n) if ob == "N" else (n, k)) for _ in range(batch)]
    Cm = [_rand(rng, (m, n)) for _ in range(batch)]
    alpha, beta = 0.7 - 0.2j, -0.3 + 0.5j
    ref = [alpha * _op(a, oa) @ _op(b, ob) + beta * c for a, b, c in zip(A, B, Cm)]
    dev = lambda mats: torch.from_numpy(np.stack([np.asfortranarray(x).ravel(order="F") for x in mats])).cuda()
    dA, dB, dC = dev(A), dev(B), dev(Cm)
    lda = A[0].shape[0]
    ldb = B[0].shape[0]
    al = np.array([alpha.real, alpha.imag])
    be = np.array([beta.real, beta.imag])
    hip.check(hip.load().hipnegf_zgemm_batched(
        eng.handle, m, n, k, al.ctypes.data, dA.data_ptr(), dA.shape[1], lda, OPS[oa],
        dB.data_ptr(), dB.shape[1], ldb, OPS[ob], be.ctypes.data, dC.data_ptr(), dC.shape[1], m, batch))
    out = dC.cpu().numpy()
    for b in range(batch):
        got = out[b].reshape((m, n), order="F")
        bound = np.abs(_op(A[b], oa)) @ np.abs(_op(B[b], ob)) + np.abs(Cm[b])
        assert np.max(np.abs(got - ref[b]) / bound) < 1e-13


@pytest.mark.parametrize("m,n,k", [(64, 64, 64), (100, 37, 129)])
def test_cgemm_batched(gpu, m, n, k):
    torch, hip, eng = gpu
    rng = np.random.default_rng(7)
    batch = 2
    A = [_rand(rng, (m, k)).astype(np.complex64) for _ in range(batch)]
    B = [_rand(rng, (k, n)).astype(np.complex64) for _ in range(batch)]
    dev = lambda mats: torch.from_numpy(np.stack([np.asfortranarray(x).ravel(order="F") for x in mats])).cuda()
    dA, dB = dev(A), dev(B)
    dC = torch.zeros((batch, m * n), dtype=torch.complex64, device="cuda")
    al, be = np.array([1.0, 0.0]), np.array([0.0, 0.0])
    hip.check(hip.load().hipnegf_cgemm_batched(
        eng.handle, m, n, k, al.ctypes.data, dA.data_ptr(), m * k, m, 0, dB.data_ptr(), k * n, k, 0,
        be.ctypes.data, dC.data_ptr(), m * n, m, batch))
    out = dC.cpu().numpy()
    for b in range(batch):
        got = out[b].reshape((m, n), order="F")
        ref = A[b].astype(np.complex128) @ B[b].astype(np.complex128)
        bound = np.abs(A[b]) @ np.abs(B[b])
        assert np.max(np.abs(got - ref) / bound) < 4e-6


def _inv_case(rng, n, kind):
    if kind == "rand":
        return _rand(rng, (n, n))
    if kind == "perm":  # zero diagonal: forces pivoting everywhere
        P = np.eye(n)[rng.permutation(n)]
        return P * (1.0 + rng.random((n, n))) + 1e-3 * _rand(rng, (n, n))
    if kind == "green":  # E - H + i eta with a random Hermitian H
        H = _rand(rng, (n, n))
        H = (H + H.conj().T) / 2
        return (0.3 + 1e-3j) * np.eye(n) - H


@pytest.mark.parametrize("n", [1, 2, 5, 10, 16, 17, 31, 32, 33, 63, 64, 65, 72, 80, 81, 95, 96, 97, 100, 101, 128,
                               130, 200, 256, 300, 383, 449, 512, 513, 700, 1024])
@pytest.mark.parametrize("kind", ["rand", "perm", "green"])
def test_zinverse_batched(gpu, n, kind):
    torch, hip, eng = gpu
    rng = np.random.default_rng(n + len(kind))
    batch = 3
    A = [_inv_case(rng, n, kind) for _ in range(batch)]
    dA = torch.from_numpy(np.stack([a.ravel(order="F") for a in A])).cuda()
    dX = torch.zeros_like(dA)
    st = torch.zeros(batch, dtype=torch.int32, device="cuda")
    hip.check(hip.load().hipnegf_zinverse_batched(eng.handle, n, dA.data_ptr(), n * n, dX.data_ptr(),
                                                  n * n, batch, st.data_ptr()))
    X = dX.cpu().numpy()
    assert st.cpu().numpy().tolist() == [0] * batch
    for b in range(batch):
        x = X[b].reshape((n, n), order="F")
        res = np.abs(A[b] @ x - np.eye(n)).max() / (np.abs(A[b]).max() * np.abs(x).max() * n)
        assert res < 1e-13, res
        ref = np.linalg.inv(A[b])
        cond = np.linalg.cond(A[b])
        assert np.abs(x - ref).max() <= 1e-14 * cond * n * np.abs(ref).max()


def test_zinverse_singular_flag(gpu):
    torch, hip, eng = gpu
    for n in (8, 160, 700):
        A = np.ones((n, n), dtype=np.complex128)
        dA = torch.from_numpy(A.ravel(order="F").copy()).cuda().reshape(1, -1)
        dX = torch.zeros_like(dA)
        st = torch.zeros(1, dtype=torch.int32, device="cuda")
        hip.check(hip.load().hipnegf_zinverse_batched(eng.handle, n, dA.data_ptr(), n * n,
                                                      dX.data_ptr(), n * n, 1, st.data_ptr()))
        assert int(st.cpu()[0]) & hip.ST_SINGULAR


@pytest.mark.parametrize("n", [10, 64, 128, 200, 256, 320, 512, 777])
def test_cinverse_batched(gpu, n):
    torch, hip, eng = gpu
    rng = np.random.default_rng(n)
    batch = 2
    A = [_inv_case(rng, n, "rand").astype(np.complex64) for _ in range(batch)]
    dA = torch.from_numpy(np.stack([a.ravel(order="F") for a in A])).cuda()
    dX = torch.zeros_like(dA)
    st = torch.zeros(batch, dtype=torch.int32, device="cuda")
    hip.check(hip.load().hipnegf_cinverse_batched(eng.handle, n, dA.data_ptr(), n * n, dX.data_ptr(),
                                                  n * n, batch, st.data_ptr()))
    X = dX.cpu().numpy()
    for b in range(batch):
        x = X[b].reshape((n, n), order="F").astype(np.complex128)
        a = A[b].astype(np.complex128)
        res = np.abs(a @ x - np.eye(n)).max() / (np.abs(a).max() * np.abs(x).max() * n)
        assert res < 1e-5, res


@pytest.mark.parametrize("m,n,k,beta,ops", [(64, 64, 64, 0.0, (0, 0)), (128, 128, 128, 1.0, (0, 0)),
                                            (256, 256, 64, 1.0, (0, 0)), (100, 200, 72, 0.0, (0, 0)),
                                            (32, 32, 32, 1.0, (0, 0)), (48, 40, 24, 1.0, (1, 2)),
                                            (70, 90, 33, 0.0, (2, 1)), (64, 64, 64, 1.0, (2, 2))])
def test_gemm_tile_variants_bitwise(gpu, m, n, k, beta, ops):
    """The autotuner's three FP64 tilings (64 x 64 four-wave, 64 x 128
    eight-wave, 32 x 32 single-wave) run the same MFMA chains in the same
    order: bitwise the same C for every operand form, so the tuner's choice
    never changes a result.  Then the tuner itself records the shape class."""
    torch, hip, eng = gpu
    rng = np.random.default_rng(m + n + k + 7 * ops[0] + ops[1])
    batch = 5
    opa, opb = ops
    ra, ca = (m, k) if opa == 0 else (k, m)   # stored shapes (column-major, ld = rows)
    rb, cb = (k, n) if opb == 0 else (n, k)
    A = rng.standard_normal((batch, ra * ca)) + 1j * rng.standard_normal((batch, ra * ca))
    B = rng.standard_normal((batch, rb * cb)) + 1j * rng.standard_normal((batch, rb * cb))
    C0 = rng.standard_normal((batch, m * n)) + 1j * rng.standard_normal((batch, m * n))
    dA, dB = torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda()
    al = np.array([1.0, 0.5])
    be = np.array([beta, 0.0])
    out = []
    try:
        for v in (0, 1, 2, -1):
            hip.gemm_force_variant(v)
            dC = torch.from_numpy(C0.copy()).cuda()
            hip.check(hip.load().hipnegf_zgemm_batched(eng.handle, m, n, k, hip._ptr(al), dA.data_ptr(),
                                                        ra * ca, ra, opa, dB.data_ptr(), rb * cb, rb, opb,
                                                        hip._ptr(be), dC.data_ptr(), m * n, m, batch))
            eng.sync()
            out.append(dC.cpu().numpy())
    finally:
        hip.gemm_force_variant(-1)
    for o in out[1:]:
        np.testing.assert_array_equal(o, out[0])
    op = lambda X, o: X if o == 0 else (X.T if o == 1 else X.conj().T)
    ref = np.stack([(al[0] + 1j * al[1]) * (op(A[b].reshape((ra, ca), order="F"), opa)
                                            @ op(B[b].reshape((rb, cb), order="F"), opb))
                    + beta * C0[b].reshape((m, n), order="F") for b in range(batch)])
    got = np.stack([out[0][b].reshape((m, n), order="F") for b in range(batch)])
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()
    st = hip.gemm_tune_stats()
    assert st["on"] and st["shapes"] >= 1
