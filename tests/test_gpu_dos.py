"""GPU parity tests of transmission + local DOS (C ABI
hipnegf_solve_transmission_dos, calculate_transmissions_and_dos in
iterative.F90:1821-1952) and of the tun_proj trace mask, against the CPU
oracle on the same inputs; the libNEGF C API layer (negf_init_ldos /
negf_set_ldos_intervals / negf_associate_ldos) on the reference's Python
example.

Bar (complex FP64): T within 1e-10 relative (absolute 1e-10 max T for tiny
values), every LDOS value within 1e-10 of the oracle relative to the largest
LDOS value of the test, identical decimation counts.
"""
import numpy as np
import pytest
import scipy.sparse as sp

from libnegf_amd import workloads as wl
from libnegf_amd.structure import BlockStructure, Contact
from tests.chains import chain_dense, from_dense

pytestmark = pytest.mark.gpu

RTOL = 1e-10


@pytest.fixture(scope="module")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from libnegf_amd import hip
    from oracle import pyoracle
    pyoracle.build()
    pyoracle.set_backend("builtin")
    return hip, hip.Engine(0), pyoracle


def _cmpT(T, Tref, rtol=RTOL):
    scale = max(np.abs(Tref).max(), 1e-300)
    big = Tref > 1e-6 * scale
    if big.any():
        assert (np.abs(T[big] - Tref[big]) / Tref[big]).max() <= rtol
    assert np.abs(T[~big] - Tref[~big]).max(initial=0) <= rtol * scale


def _cmpL(L, Lref, rtol=RTOL):
    assert L.shape == Lref.shape
    assert np.abs(L - Lref).max() <= rtol * max(np.abs(Lref).max(), 1e-300)


def _oracle_grid(orc, s, E, delta, dd, pairs, proj=None, full=False, tun=None, eps=0.0):
    rs = [orc.transmission_dos(s, e, delta, dd, pairs=pairs, proj=proj, full=full, tun=tun,
                               eps_drop=eps) for e in E]
    return (np.array([r["T"] for r in rs]), np.array([r["ldos"] for r in rs]),
            np.array([r["ncyc"] for r in rs]))


def _kslab_off_block(n=24, nbl=6, seed=5):
    base = wl.KSlab.make(n=n, nbl=nbl, seed=seed).at_k(0.2, 0.9)
    c0, c1 = base.contacts
    mk = lambda c, b: Contact(cblk=b, H00=c.H00, H01=c.H01, HMC=c.HMC, S00=c.S00, S01=c.S01, SMC=c.SMC)
    return BlockStructure(Hd=base.Hd, Hu=base.Hu, Hl=base.Hl, Sd=base.Sd, Su=base.Su, Sl=base.Sl,
                          contacts=[mk(c0, nbl - 1), mk(c1, 2)])


@pytest.mark.parametrize("dd", [0.0, 1e-3])
def test_kslab_two_contacts(gpu, dd):
    """Overlap slab, 2 contacts (contact 2 on block 1): T of both pair orders,
    per-orbital DOS and interval projections."""
    hip, eng, orc = gpu
    s = wl.KSlab.make(n=32, nbl=5, seed=3).at_k(0.7, -0.4)
    E = np.linspace(-1.2, 1.3, 7)
    ds = hip.DeviceStructure(s)
    pairs = [(1, 2), (2, 1)]
    T, L, ncyc, st = hip.solve_transmission_dos(eng, ds, E, pairs=pairs, full=True, dos_delta=dd,
                                                delta=1e-5)
    Tr, Lr, nr = _oracle_grid(orc, s, E, 1e-5, dd, pairs, full=True)
    assert (st == 0).all()
    _cmpT(T[:, 0], Tr[:, 0])
    _cmpT(T[:, 1], Tr[:, 1])
    _cmpL(L, Lr)
    assert (ncyc == nr).all()
    proj = [list(range(1, 33)), [5, 5, 70], list(range(100, 161)), [s.ndev, s.ndev + 3]]
    T2, L2, _, _ = hip.solve_transmission_dos(eng, ds, E, pairs=pairs, proj=proj, dos_delta=dd,
                                              delta=1e-5)
    _, Lr2, _ = _oracle_grid(orc, s, E, 1e-5, dd, pairs, proj=proj)
    np.testing.assert_array_equal(T2, T)
    _cmpL(L2, Lr2)


def test_off_first_block_and_nonuniform(gpu):
    """Lower contact on block 2 (2-contact formula on Gr(2,2)); and a device
    with unequal block sizes from a dense H."""
    hip, eng, orc = gpu
    s = _kslab_off_block()
    E = np.array([-0.8, 0.1, 0.9])
    ds = hip.DeviceStructure(s)
    T, L, ncyc, st = hip.solve_transmission_dos(eng, ds, E, pairs=[(1, 2)], full=True, dos_delta=5e-4,
                                                delta=1e-5)
    Tr, Lr, nr = _oracle_grid(orc, s, E, 1e-5, 5e-4, [(1, 2)], full=True)
    assert (st == 0).all()
    _cmpT(T[:, 0], Tr[:, 0])
    _cmpL(L, Lr)
    assert (ncyc == nr).all()
    # item groups on concurrent streams: bitwise the same
    T3, L3, n3, s3 = hip.solve_transmission_dos(eng, ds, E, pairs=[(1, 2)], full=True, dos_delta=5e-4,
                                                delta=1e-5, lanes=3)
    np.testing.assert_array_equal(T3, T)
    np.testing.assert_array_equal(L3, L)
    np.testing.assert_array_equal(n3, ncyc)
    # unequal PLs: 100-site chain with disorder, PLs of 9, 17, 11, 23 sites
    rng = np.random.default_rng(8)
    H = chain_dense(onsite=0.4 * rng.standard_normal(60))
    plend = [9, 26, 37, 60]
    s2 = from_dense(H, None, plend, [61, 81], [80, 100], [4, 1])
    ds2 = hip.DeviceStructure(s2)
    E2 = np.linspace(-1.9, 1.9, 9)
    proj = [[1, 9, 10, 26], list(range(27, 61))]
    T, L, _, st = hip.solve_transmission_dos(eng, ds2, E2, proj=proj, dos_delta=1e-4, delta=1e-6)
    Tr, Lr, _ = _oracle_grid(orc, s2, E2, 1e-6, 1e-4, [(1, 2)], proj=proj)
    assert (st == 0).all()
    _cmpT(T[:, 0], Tr[:, 0])
    _cmpL(L, Lr)


def test_three_contacts(gpu):
    """N-contact column sweep for every pair.  delta = 1e-4 (the reference's
    default): at 1e-5 this ribbon's contact blocks are conditioned so that the
    oracle's own built-in and MKL backends differ by up to 2e-9 in T."""
    hip, eng, orc = gpu
    base = wl.ribbon(10, 6, disorder=0.4, seed=31)
    c0 = base.contacts[0]
    side = Contact(cblk=3, H00=c0.H00, H01=c0.H01, HMC=-0.6 * np.eye(10))
    s = BlockStructure(Hd=base.Hd, Hu=base.Hu, Hl=base.Hl, contacts=[c0, base.contacts[1], side])
    E = np.linspace(-2.4, 2.4, 7)
    pairs = [(1, 2), (1, 3), (2, 3), (3, 1)]
    ds = hip.DeviceStructure(s)
    proj = [list(range(1, 11)), list(range(21, 31))]
    T, L, ncyc, st = hip.solve_transmission_dos(eng, ds, E, pairs=pairs, proj=proj, dos_delta=2e-4,
                                                delta=1e-4)
    Tr, Lr, nr = _oracle_grid(orc, s, E, 1e-4, 2e-4, pairs, proj=proj)
    assert (st == 0).all()
    for j in range(len(pairs)):
        _cmpT(T[:, j], Tr[:, j])
    _cmpL(L, Lr)
    assert (ncyc == nr).all()
    # with the same pairs, T equals the multi-terminal transmission at dos_delta 0
    T0, _, _, _ = hip.solve_transmission_dos(eng, ds, E, pairs=pairs, proj=proj, delta=1e-4)
    Tm, _, _ = hip.solve_transmission_pairs(eng, ds, E, pairs, delta=1e-4)
    np.testing.assert_allclose(T0, Tm, rtol=1e-11, atol=1e-13)


def test_one_contact_dos_only(gpu):
    hip, eng, orc = gpu
    base = wl.ribbon(8, 5, disorder=0.3, seed=2)
    s = BlockStructure(Hd=base.Hd, Hu=base.Hu, Hl=base.Hl, contacts=[base.contacts[0]])
    E = np.array([-0.7, 0.4])   # E = -1 is a subband edge of the W = 8 lead (singular g_s)
    ds = hip.DeviceStructure(s)
    T, L, _, st = hip.solve_transmission_dos(eng, ds, E, pairs=[(1, 2)], full=True, delta=1e-5)
    Tr, Lr, _ = _oracle_grid(orc, s, E, 1e-5, 0.0, [(1, 2)], full=True)
    assert (st == 0).all() and (T == 0).all()
    _cmpL(L, Lr)


def test_batch_split_offsets(gpu):
    hip, eng, orc = gpu
    s = wl.KSlab.make(n=16, nbl=4, seed=9).at_k(-0.3, 0.5)
    E = np.linspace(-1.0, 1.0, 11)
    ds = hip.DeviceStructure(s)
    proj = [[1, 2], [17, 40, 64], list(range(1, 65))]
    a = hip.solve_transmission_dos(eng, ds, E, proj=proj, dos_delta=1e-3, delta=1e-5)
    b = hip.solve_transmission_dos(eng, ds, E, proj=proj, dos_delta=1e-3, delta=1e-5, max_batch=3)
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)


def test_tun_proj_mask(gpu):
    """tun_proj rows on the fast path (block 1), the LDOS path and the
    N-contact path (block of the deeper contact)."""
    hip, eng, orc = gpu
    s = wl.KSlab.make(n=20, nbl=4, seed=5).at_k(0.1, 0.2)
    E = np.array([-0.6, 0.4])
    ds = hip.DeviceStructure(s)
    T_all, _, _ = hip.solve_transmission(eng, ds, E, delta=1e-5)
    tun = [1, 4, 5, 9, 17, 33, 70]           # rows 1..20 are block 1
    ds.set_tun_proj(tun)
    T, _, st = hip.solve_transmission(eng, ds, E, delta=1e-5)
    Td, _, _, _ = hip.solve_transmission_dos(eng, ds, E, delta=1e-5)
    ref = np.array([orc.transmission_dos(s, e, 1e-5, 0.0, tun=tun)["T"][0] for e in E])
    assert (st == 0).all()
    _cmpT(T, ref)
    _cmpT(Td[:, 0], ref)
    assert np.abs(T - T_all).max() > 1e-6 * T_all.max()
    ds.set_tun_proj(None)
    T2, _, _ = hip.solve_transmission(eng, ds, E, delta=1e-5)
    np.testing.assert_array_equal(T2, T_all)
    # N-contact: mask rows of the deeper contact's block
    base = wl.ribbon(8, 5, disorder=0.4, seed=6)
    c0 = base.contacts[0]
    side = Contact(cblk=3, H00=c0.H00, H01=c0.H01, HMC=-0.5 * np.eye(8))
    s3 = BlockStructure(Hd=base.Hd, Hu=base.Hu, Hl=base.Hl, contacts=[c0, base.contacts[1], side])
    ds3 = hip.DeviceStructure(s3)
    tun3 = [17, 19, 20, 24, 33, 40, 2]
    ds3.set_tun_proj(tun3)
    T3, _, _ = hip.solve_transmission_pairs(eng, ds3, E, [(2, 3), (1, 3)], delta=1e-5)
    for j, (a, b) in enumerate([(2, 3), (1, 3)]):
        ref3 = []
        for e in E:
            sig = [orc.contact_sigma(c, s3.bsz[c.cblk - 1], e + 1j * 1e-5)[0] for c in s3.contacts]
            ref3.append(orc.transmission_sigma_tun(s3, e, sig, a, b, tun=tun3))
        _cmpT(T3[:, j], np.array(ref3))


# ------------------------------------------------------------ C API layer ----
def _chain_session(NEGF, H, plend, cblk, emin, emax, estep):
    n = NEGF()
    n.set_h(sp.csr_matrix(H))
    n.set_s_id(H.shape[0])
    n.init_structure(2, np.array([80, 100]), np.array([60, 80]), len(plend), np.array(plend),
                     np.array(cblk), surfstart=np.array([61, 81]))
    p = n.params
    p.emin, p.emax, p.estep = emin, emax, estep
    return n


def test_api_hello_ctypes_ldos(gpu):
    """tests/cpython/hello_ctypes.py: double-barrier chain, LDOS on orbitals
    1-60, 31-60, 1-30, then solve_landauer; T, LDOS (times kwght) and the
    current vs the oracle."""
    hip, eng, orc = gpu
    from libnegf_amd.negf import NEGF
    ons = np.zeros(60)
    ons[25] = ons[35] = 1.0
    H = chain_dense(onsite=ons)
    plend = [15, 30, 45, 60]
    n = _chain_session(NEGF, H, plend, [4, 1], -3.0, 3.0, 0.05)
    n.params.mu[0] = 0.1
    n.params.kwght = 0.5
    n.params.dos_delta = 1e-4
    n.set_ldos_intervals(np.array([1, 31, 1]), np.array([60, 60, 30]))
    n.solve_landauer()
    E, _ = n.get_energies()
    T = n.get_transmission()
    L = n.get_ldos()
    assert T.shape == (1, len(E)) and L.shape == (3, len(E))
    s = from_dense(H, None, plend, [61, 81], [80, 100], [4, 1])
    proj = [list(range(1, 61)), list(range(31, 61)), list(range(1, 31))]
    Tr, Lr, _ = _oracle_grid(orc, s, E, 1e-4, 1e-4, [(1, 2)], proj=proj)
    _cmpT(T[0], 0.5 * Tr[:, 0])
    _cmpL(L.T, 0.5 * Lr)
    n.close()


def test_api_full_orbital_ldos_and_indexes(gpu):
    """One projection per orbital with dos_proj(1)%indexes(1) == 0 gives the
    per-orbital DOS (iterative.F90:1930-1931); set_ldos_indexes lists; the
    ideal chain's LDOS is 1 / (pi sqrt(4 - E^2)) per site."""
    hip, eng, orc = gpu
    from libnegf_amd.negf import NEGF
    H = chain_dense()
    n = _chain_session(NEGF, H, [10, 20, 30, 40, 50, 60], [6, 1], -1.5, 1.5, 0.25)
    n.params.delta = 1e-9
    n.init_ldos(60)
    n.set_ldos_indexes(1, [0])
    for k in range(2, 61):
        n.set_ldos_indexes(k, [k])
    n.solve_landauer()
    E, _ = n.get_energies()
    L = n.get_ldos()
    assert L.shape == (60, len(E))
    rho = 1.0 / (np.pi * np.sqrt(4.0 - E ** 2))
    assert np.abs(L / rho[None, :] - 1.0).max() <= 1e-6
    # projections by explicit index lists
    n.init_ldos(2)
    n.set_ldos_indexes(1, [1, 2, 3, 99])
    n.set_ldos_indexes(2, [60, 60])
    n.solve_landauer()
    L2 = n.get_ldos()
    np.testing.assert_allclose(L2[0], 3 * rho, rtol=1e-6)
    np.testing.assert_allclose(L2[1], 2 * rho, rtol=1e-6)
    # tun_proj through the C API: all rows of block 1 -> unchanged T
    T_all = n.get_transmission().copy()
    n.set_tun_indexes(list(range(1, 11)))
    n.solve_landauer()
    np.testing.assert_array_equal(n.get_transmission(), T_all)
    n.set_tun_indexes([])
    n.solve_landauer()
    assert (n.get_transmission() == 0).all()
    n.close()
