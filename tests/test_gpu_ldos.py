"""compute_ldos (libnegf.F90:1893-1901; ldos_int, integrations.F90:212-258)
through the C API (negf_compute_ldos): ldos_mat(i, l) = sum over the
orbitals of projection l of -g_spin Im diag(G^r S) / pi at E_i + i dos_delta,
with G^r's device/contact blocks (compute_Gr with outer = 1) and S the
whole overlap, device and contacts.

References:
  * the oracle's G^r blocks at the same complex energy (gr_point: device and
    contacts at E + i dos_delta) multiplied with the structure's S blocks;
  * independent of the block recursion: a dense inverse of the whole device
    matrix z S - H - sum Sigma_k with the oracle's contact self-energies, and
    G^r_DC = -G^r Tlc GS from its surface functions;
  * with S = I and delta = dos_delta, g_spin / kwght times the LDOS that
    solve_landauer returns (calculate_transmissions_and_dos), and the ideal
    chain's closed form g_spin / (pi sqrt(4 - E^2)) per site.
Bar: within 1e-10 of the largest |LDOS| of the test (FP64).
"""
import numpy as np
import pytest
import scipy.sparse as sp

from tests.chains import chain_dense, from_dense

pytestmark = pytest.mark.gpu

RTOL = 1e-10
PLEND = [15, 30, 45, 60]
CBLK = [4, 1]


@pytest.fixture(scope="module")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from libnegf_amd import hip
    from libnegf_amd.negf import NEGF
    from oracle import pyoracle
    pyoracle.build()
    pyoracle.set_backend("builtin")
    return hip, NEGF, pyoracle


def _overlap_chain(seed=4, s_nn=0.12):
    """The reference tests' 100-site chain with on-site disorder and a
    nearest-neighbour overlap on every bond of H (device and contacts)."""
    rng = np.random.default_rng(seed)
    H = chain_dense(onsite=0.3 * rng.standard_normal(60))
    A = (np.abs(H) > 0).astype(float)
    np.fill_diagonal(A, 0.0)
    S = np.eye(H.shape[0]) + s_nn * A
    return H, S


def _session(NEGF, H, S, emin, emax, estep):
    n = NEGF()
    n.set_h(sp.csr_matrix(H))
    if S is None:
        n.set_s_id(H.shape[0])
    else:
        n.set_s(sp.csr_matrix(S))
    n.init_structure(2, np.array([80, 100]), np.array([60, 80]), len(PLEND), np.array(PLEND),
                     np.array(CBLK), surfstart=np.array([61, 81]))
    p = n.params
    p.emin, p.emax, p.estep = emin, emax, estep
    return n


def _diag_blocks(orc, s, z, outer=True):
    """diag(G^r S) over the device from the oracle's G^r blocks (outer: with
    the device/contact terms G^r_DC S_CD)."""
    r = orc.gr_point(s, complex(z))
    nbl = len(s.Hd)
    eye = lambda m: np.eye(m.shape[0])
    out = []
    for i in range(nbl):
        Sd = s.Sd[i] if s.Sd is not None else eye(s.Hd[i])
        d = np.diag(r["Gd"][i] @ Sd).copy()
        if s.Su is not None:
            if i > 0:
                d += np.diag(r["Gl"][i - 1] @ s.Su[i - 1])
            if i < nbl - 1:
                d += np.diag(r["Gu"][i] @ s.Sl[i])
        for k, c in enumerate(s.contacts):
            if outer and c.cblk - 1 == i and c.SMC is not None:
                assert r["Glc"][k].shape == c.SMC.shape
                d += np.diag(r["Glc"][k] @ c.SMC.conj().T)
        out.append(d)
    return np.concatenate(out)


def _diag_dense(orc, s, H, S, z):
    """The same from a dense inverse of the whole device (no block recursion)."""
    nd = sum(h.shape[0] for h in s.Hd)
    off = np.cumsum([0] + [h.shape[0] for h in s.Hd])
    Sdd = np.eye(nd) if S is None else S[:nd, :nd]
    A = z * Sdd - H[:nd, :nd]
    outer = []
    for c in s.contacts:
        b = c.cblk - 1
        Sig, GS, _ = orc.contact_sigma(c, s.Hd[b].shape[0], complex(z))
        A[off[b]:off[b + 1], off[b]:off[b + 1]] -= Sig
        outer.append((b, c, GS))
    G = np.linalg.inv(A)
    d = np.diag(G @ Sdd).copy()
    for b, c, GS in outer:
        if c.SMC is None:
            continue
        Tlc = z * c.SMC - c.HMC
        Gdc = -G[off[b]:off[b + 1], off[b]:off[b + 1]] @ Tlc @ GS
        d[off[b]:off[b + 1]] += np.diag(Gdc @ c.SMC.conj().T)
    return d


def _project(diag, proj, nd, g_spin=2.0):
    v = -g_spin * diag.imag / np.pi
    return np.array([sum(v[i - 1] for i in p if i <= nd) for p in proj])


def test_compute_ldos_overlap_chain(gpu):
    """Overlap chain, both contacts with S_DC != 0: every projection against
    the oracle's blocks and the dense inverse; intervals that run past the
    device into a contact count only the device orbitals."""
    hip, NEGF, orc = gpu
    H, S = _overlap_chain()
    n = _session(NEGF, H, S, -1.7, 1.9, 0.3)
    n.params.dos_delta = 1e-3
    istart, iend = np.array([1, 31, 1, 55, 14]), np.array([60, 60, 30, 70, 16])
    n.set_ldos_intervals(istart, iend)
    n.compute_ldos()
    E, _ = n.get_energies()
    L = n.get_ldos()
    n.close()
    assert L.shape == (len(istart), len(E))
    s = from_dense(H, S, PLEND, [61, 81], [80, 100], CBLK)
    proj = [list(range(a, b + 1)) for a, b in zip(istart, iend)]
    ref_b = np.array([_project(_diag_blocks(orc, s, e + 1e-3j), proj, 60) for e in E])
    ref_d = np.array([_project(_diag_dense(orc, s, H, S, e + 1e-3j), proj, 60) for e in E])
    scale = np.abs(ref_b).max()
    assert np.abs(ref_b - ref_d).max() <= 1e-11 * scale   # the two references agree
    assert np.abs(L.T - ref_b).max() <= RTOL * scale
    # the device/contact terms matter here: without them the intervals on
    # the contact blocks (1-30 holds block 1, 55-70 block 4) move well past the bar
    ref_in = np.array([_project(_diag_blocks(orc, s, e + 1e-3j, outer=False), proj, 60) for e in E])
    assert np.abs(ref_in[:, 2] - ref_b[:, 2]).max() > 1e-4 * scale
    assert np.abs(ref_in[:, 3] - ref_b[:, 3]).max() > 1e-4 * scale
    np.testing.assert_array_equal(ref_in[:, 4], ref_b[:, 4])  # 14-16: no contact block


def test_compute_ldos_identity_s(gpu):
    """S = I: with delta = dos_delta the LDOS is g_spin / kwght times
    solve_landauer's (both -Im diag G^r / pi at the same point); on the ideal
    chain each site's value is g_spin / (pi sqrt(4 - E^2)) up to O(dos_delta)."""
    hip, NEGF, orc = gpu
    ons = np.zeros(60)
    ons[25] = ons[35] = 1.0
    H = chain_dense(onsite=ons)
    istart, iend = np.array([1, 31, 1, 26]), np.array([60, 60, 30, 26])
    n = _session(NEGF, H, None, -1.5, 1.5, 0.1)
    n.params.dos_delta = 1e-4
    n.params.delta = 1e-4
    n.params.kwght = 0.5
    n.set_ldos_intervals(istart, iend)
    n.compute_ldos()
    L = n.get_ldos().copy()
    n.solve_landauer()
    Lt = n.get_ldos().copy()
    E, _ = n.get_energies()
    n.close()
    scale = np.abs(L).max()
    assert np.abs(L - (2.0 / 0.5) * Lt).max() <= 1e-12 * scale
    s = from_dense(H, None, PLEND, [61, 81], [80, 100], CBLK)
    proj = [list(range(a, b + 1)) for a, b in zip(istart, iend)]
    ref = np.array([_project(_diag_blocks(orc, s, e + 1e-4j), proj, 60) for e in E])
    assert np.abs(L.T - ref).max() <= RTOL * scale
    # ideal chain: one site, closed form (g_spin = 2)
    n = _session(NEGF, chain_dense(), None, -1.5, 1.5, 0.25)
    n.params.dos_delta = 1e-9
    n.set_ldos_intervals(np.array([30]), np.array([30]))
    n.compute_ldos()
    E, _ = n.get_energies()
    L = n.get_ldos()[0]
    n.close()
    np.testing.assert_allclose(L, 2.0 / (np.pi * np.sqrt(4.0 - E ** 2)), rtol=1e-6)


def test_compute_ldos_rejects(gpu):
    hip, NEGF, orc = gpu
    H, S = _overlap_chain()
    n = _session(NEGF, H, S, -1.0, 1.0, 0.5)
    with pytest.raises(Exception, match="projections"):
        n.compute_ldos()
    n.set_ldos_intervals(np.array([1]), np.array([101]))
    with pytest.raises(Exception, match="outside"):
        n.compute_ldos()
    n.close()
