"""negf_density_efa on the GPU against the oracle (compute_density_efa,
libnegf.F90:1690-1791): electron and hole densities from the contour, the
real-axis trapezoid, or both, under bias and with an overlap; q and the whole
density matrix (pattern and values) at 1e-10 of the largest entry.  Plus the
ideal chain's sum rule n + p = 2 per site."""
import numpy as np
import pytest

from tests.chains import chain_dense, from_dense
from tests.test_gpu_density import RTOL, _cmp_dm, _session, gpu  # noqa: F401  (fixture)

pytestmark = pytest.mark.gpu


def _params(p, nc):
    g = lambda a: [a[k] for k in range(nc)]
    return dict(mu=g(p.mu), mu_n=g(p.mu_n), mu_p=g(p.mu_p), min_or_max=p.min_or_max,
                np_n=[p.np_n[0], p.np_n[1]], np_p=[p.np_p[0], p.np_p[1]], n_poles=p.n_poles, ec=p.ec,
                deltaec=p.deltaec, ev=p.ev, deltaev=p.deltaev, kbt_dm=g(p.kbt_dm), n_kt=p.n_kt,
                g_spin=p.g_spin, np_real=p.np_real, delta=p.delta, dore=bytes(p.dore)[:1].decode())


@pytest.mark.parametrize("particle", [1, -1])
@pytest.mark.parametrize("contour,real,dore", [(True, True, "D"), (True, False, "E"), (False, True, "D")])
def test_efa_overlap_bias(gpu, particle, contour, real, dore):
    hip, eng, orc = gpu
    rng = np.random.default_rng(7)
    H = chain_dense(onsite=0.3 * rng.standard_normal(60))
    S = np.eye(100, dtype=np.complex128)
    for i in range(1, 100):
        if H[i, i - 1] != 0:
            S[i, i - 1] = S[i - 1, i] = 0.07
    plend = [12, 25, 40, 60]
    n = _session(H, S, plend, [4, 1], [61, 81], [80, 100])
    p = n.params
    p.ec, p.deltaec, p.ev, p.deltaev = -3.2, 0.1, 3.1, 0.2
    p.mu[0], p.mu[1] = 0.25, -0.1
    p.mu_n[0], p.mu_n[1] = 0.2, -0.05
    p.mu_p[0], p.mu_p[1] = 0.15, -0.2
    p.kbt_dm[0], p.kbt_dm[1] = 0.02, 0.01
    npts = (12, 14) if contour else (0, 0)
    p.np_n[0], p.np_n[1] = npts
    p.np_p[0], p.np_p[1] = npts
    p.n_poles = 3 if contour else 0
    p.np_real = 17 if real else 0
    p.dore = dore.encode()
    q = n.density_efa(100, particle)
    dm = n.get_dm()
    s = from_dense(H, S, plend, [61, 81], [80, 100], [4, 1])
    rho, pat, qref = orc.density_efa(s, 100, [60, 80], S != 0, particle, _params(n.params, 2))
    assert np.abs(q - qref).max() <= RTOL * np.abs(qref).max()
    _cmp_dm(dm, rho, pat)
    n.close()


def test_efa_chain_sum_rule(gpu):
    hip, eng, orc = gpu
    H = chain_dense()
    qs = []
    for particle in (1, -1):
        n = _session(H, None, [15, 30, 45, 60], [4, 1], [61, 81], [80, 100])
        p = n.params
        p.ec, p.ev = -3.0, 3.0
        for k in range(2):
            p.mu[k] = p.mu_n[k] = p.mu_p[k] = 0.3
            p.kbt_dm[k] = 0.05
        p.np_n[0], p.np_n[1] = 60, 150
        p.np_p[0], p.np_p[1] = 60, 150
        p.n_poles, p.n_kt, p.np_real = 4, 30, 0
        qs.append(n.density_efa(60, particle))
        n.close()
    assert np.abs(qs[0] + qs[1] - 2.0).max() < 1e-8


def test_efa_errors(gpu):
    from libnegf_amd.negf import NegfError
    n = _session(chain_dense(), None, [15, 30, 45, 60], [4, 1], [61, 81], [80, 100])
    with pytest.raises(NegfError, match="unknown particle"):
        n.density_efa(60, 0)
    with pytest.raises(NegfError, match="bulk contact density"):
        n.density_efa(60, 1, contacts=True)
    n.close()


@pytest.mark.parametrize("particle", [1, -1])
def test_quasi_equilibrium_profiles(gpu, particle):
    """negf_density_quasi_equilibrium against the oracle (quasiEq_int_n/p):
    overlap chain, band-edge profiles, three quasi-Fermi levels, kwght 0.5."""
    hip, eng, orc = gpu
    rng = np.random.default_rng(11)
    H = chain_dense(onsite=0.3 * rng.standard_normal(60))
    S = np.eye(100, dtype=np.complex128)
    for i in range(1, 100):
        if H[i, i - 1] != 0:
            S[i, i - 1] = S[i - 1, i] = 0.07
    plend = [12, 25, 40, 60]
    n = _session(H, S, plend, [4, 1], [61, 81], [80, 100])
    p = n.params
    p.deltaec, p.deltaev = 1.5, 1.2
    p.kbt_dm[0], p.kbt_dm[1] = 0.02, 0.03
    p.np_n[0], p.np_n[1] = 10, 24
    p.np_p[0], p.np_p[1] = 9, 22
    p.n_poles, p.n_kt, p.kwght = 3, 10, 0.5
    nz = 70  # past the device: contact rows have no diagonal entry
    Ec = 1.0 + 0.2 * rng.standard_normal(nz)
    Ev = -1.0 + 0.2 * rng.standard_normal(nz)
    lev = np.choose(np.arange(nz) % 3, [-0.3, 0.1, 0.4]).astype(float)
    q = n.density_quasi_eq(particle, Ec, Ev, lev, lev)
    s = from_dense(H, S, plend, [61, 81], [80, 100], [4, 1])
    qref = orc.density_quasi(s, particle, dict(_params(n.params, 2), kwght=0.5), Ec, Ev, lev, lev)
    assert np.abs(q - qref).max() <= RTOL * np.abs(qref).max()
    assert np.all(q[60:] == 0.0)
    n.close()
