"""The fused small-block transmission (one launch per call, one wave per (E,k)
item running both decimations, the sweep and the trace; hn_fused.hip) against
the oracle and against the batched block-operation path (HIPNEGF_NO_FUSED)."""
import numpy as np
import pytest

from libnegf_amd import workloads as wl

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    from libnegf_amd import hip
    return hip, hip.Engine(0)


def _oracle(s, E, eta, ni=1, nf=2):
    from oracle import pyoracle
    pyoracle.build()
    pyoracle.set_backend("builtin")
    T, ncyc, _ = pyoracle.transmission_batch(s, np.asarray(E), eta, ni=ni, nf=nf)
    return T, ncyc


def _rel(T, Tref):
    big = Tref > 1e-6 * Tref.max()
    return float((np.abs(T[big] - Tref[big]) / Tref[big]).max())


def test_c1_full_grid(gpu):
    """C1 (W = 10 ribbon, 20 PLs, 200 energies): T at 1e-10 and the
    decimation counts against the oracle; the integer staircase; the same
    values as the batched path to rounding."""
    hip, eng = gpu
    s, E, eta = wl.config1()
    ds = hip.DeviceStructure(s)
    T, ncyc, st = hip.solve_transmission(eng, ds, E, delta=eta)
    assert (st == 0).all()
    Tref, nref = _oracle(s, E, eta)
    assert _rel(T, Tref) <= 1e-10
    assert np.array_equal(ncyc, nref)
    Tb, nb, _ = hip.solve_transmission(eng, ds, E, delta=eta, flags=hip.NO_FUSED)
    assert np.array_equal(nb, ncyc)
    assert _rel(T, Tb) <= 1e-11
    Ta = wl.analytic_ribbon_T(10, E)
    eps = -2 * np.cos(np.arange(1, 11) * np.pi / 11)
    edges = np.concatenate([eps - 2, eps + 2])
    far = np.array([np.min(np.abs(e - edges)) > 1e-2 for e in E])
    assert np.abs(T[far] - Ta[far]).max() < 1e-5


@pytest.mark.parametrize("n,nbl,overlap", [(6, 5, True), (16, 7, True), (13, 1, False), (9, 3, False)])
def test_slabs_against_oracle(gpu, n, nbl, overlap):
    """Random k-slabs (S != I with overlap), one PL (both contacts on block 1),
    several k-points in slots, both contact orders of the pair."""
    hip, eng = gpu
    slab = wl.KSlab.make(n=n, nbl=nbl, seed=40 + n, with_overlap=overlap)
    structs = [slab.at_k(0.3, -1.2), slab.at_k(2.0, 0.7)]
    ds = hip.DeviceStructure(structs)
    E = np.linspace(-1.5, 1.5, 9)
    EE = np.concatenate([E, E])
    sl = np.repeat([0, 1], len(E)).astype(np.int32)
    for ni, nf in ((1, 2), (2, 1)):
        T, ncyc, st = hip.solve_transmission(eng, ds, EE, slot=sl, delta=1e-5, ni=ni, nf=nf)
        assert (st == 0).all()
        for j, s in enumerate(structs):
            Tref, nref = _oracle(s, E, 1e-5, ni, nf)
            part = slice(j * len(E), (j + 1) * len(E))
            assert _rel(T[part], Tref) <= 1e-10, (ni, nf, j)
            assert np.array_equal(ncyc[part], nref)


def test_disordered_ribbon_fused_vs_batched(gpu):
    """Disordered W = 12 ribbon at eta = 1e-6 (near-singular blocks at some
    energies): the fused and batched paths take the same decimation counts as
    the oracle, and the fused T is within 1e-10 of the oracle or, where the
    batched path itself sits at its FP64 floor, within 4x the batched path's
    own distance from the oracle."""
    hip, eng = gpu
    s = wl.ribbon(12, 9, disorder=0.6, seed=3)
    E = np.linspace(-3.2, 3.2, 33)
    ds = hip.DeviceStructure(s)
    T, ncyc, st = hip.solve_transmission(eng, ds, E, delta=1e-6)
    Tb, nb, stb = hip.solve_transmission(eng, ds, E, delta=1e-6, flags=hip.NO_FUSED)
    assert (st == 0).all() and (stb == 0).all()
    assert np.array_equal(ncyc, nb)
    Tref, nref = _oracle(s, E, 1e-6)
    assert np.array_equal(ncyc, nref)
    sp = np.abs(Tb - Tref) / np.maximum(Tref, 1e-300)
    bar = np.maximum(1e-10, 4 * sp)
    big = Tref > 1e-6 * Tref.max()
    assert ((np.abs(T - Tref) / Tref)[big] <= bar[big] + 1e-10).all()
