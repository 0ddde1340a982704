"""Inelastic electron-phonon SCBA on the GPU (hipnegf_solve_inelastic_mw:
meir_wingreen with ElPhonPolarOptical / ElPhonNonPolarOptical,
integrations.F90:1804-1966, elphinel.F90, self_energy.F90) against the oracle's
restatement (or_inel_meir_wingreen) on the same inputs.

Bar (complex FP64): curr_mat and the integrated currents within 1e-10 of the
largest value; outer pass counts and the inner (elastic) pass counts equal.
"""
import numpy as np
import pytest

from tests.inel_model import LATT, bose, fermi, k_system

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


def _case(nbl, n, nk, ne, shift, kw, model="po", coupling=3.0, Emin=-1.5, Emax=1.5):
    structs, kfrac = k_system(nbl=nbl, n=n, nk=nk)
    E = np.linspace(Emin, Emax, ne)
    wq = shift * (E[1] - E[0])
    ndev = nbl * n
    iz = np.arange(ndev) // max(1, n // 2)      # two orbitals per z bin for n >= 4
    ndz = int(iz.max()) + 1
    from oracle import pyoracle
    if model == "po":
        K = pyoracle.inel_kmat_po(ndz, 1.0, kfrac, kw, coupling, wq, 1.0, 2.0, 1.0, 0.5, latt=LATT)
    else:
        K = pyoracle.inel_kmat_nonpo(ndz, 1.0, kfrac, kw, coupling * 0.05, 2.0, latt=LATT)
    frm = np.stack([fermi(E, 0.3, 0.05), fermi(E, -0.3, 0.05), np.zeros_like(E)], axis=1)
    return structs, E, K, iz, bose(wq, 0.2), frm


def _compare(hip, eng, orc, structs, kw, E, K, iz, shift, Nq, frm, tridiag, niter=40, tol=1e-9,
             coupling=None, niter_ela=0, tol_ela=1e-7, max_batch=0, integration=0):
    ds = hip.DeviceStructure(structs)
    g = hip.solve_inelastic_mw(eng, ds, E, kw, frm, 3, K, iz, shift, Nq, tridiag, niter, tol,
                               coupling=coupling, niter_ela=niter_ela, tol_ela=tol_ela,
                               integration=integration, max_batch=max_batch)
    r = orc.inelastic_mw(structs, kw, E, 1e-4, frm, 3, K, iz, shift, Nq, tridiag, niter, tol,
                         coupling=coupling, niter_ela=niter_ela, tol_ela=tol_ela,
                         integration=integration)
    ds.close()
    assert (g["status"] == 0).all()
    assert g["passes"] == r["passes"]
    np.testing.assert_array_equal(g["iters"], r["iters_ela"])
    scale = np.abs(r["curr_mat"]).max()
    assert np.abs(g["curr_mat"] - r["curr_mat"]).max() <= RTOL * scale
    assert np.abs(g["currents"] - r["currents"]).max() <= RTOL * np.abs(r["currents"]).max()
    return g, r


@pytest.mark.parametrize("tridiag", [True, False])
def test_polar_optical_vs_oracle(gpu, tridiag):
    hip, eng, orc = gpu
    kw = np.array([0.3, 0.7])
    structs, E, K, iz, Nq, frm = _case(3, 2, 2, 25, 2, kw)
    g, r = _compare(hip, eng, orc, structs, kw, E, K, iz, 2, Nq, frm, tridiag)
    assert 4 <= g["passes"] < 40


def test_nonpolar_with_elastic_dephasing(gpu):
    """Both models at once (elphdd inside each point's inner loop, the
    non-polar optical phonons in the outer loop), Simpson 1/3 currents."""
    hip, eng, orc = gpu
    kw = np.array([0.5, 0.5])
    structs, E, K, iz, Nq, frm = _case(4, 4, 2, 21, 3, kw, model="nonpo")
    cp = np.full(16, 0.15)
    g, r = _compare(hip, eng, orc, structs, kw, E, K, iz, 3, Nq, frm, True, coupling=cp, niter_ela=8,
                    tol_ela=1e-9, integration=1)
    assert g["iters"].max() > 2


def test_chunked_items_and_three_kpoints(gpu):
    """More (E,k) items than one batch (max_batch 16): the block stores are
    written and read at every batch offset."""
    hip, eng, orc = gpu
    kw = np.array([0.2, 0.3, 0.5])
    structs, E, K, iz, Nq, frm = _case(5, 4, 3, 30, 2, kw, Emin=-2.2, Emax=2.2)
    _compare(hip, eng, orc, structs, kw, E, K, iz, 2, Nq, frm, True, max_batch=16)


def test_ribbon_stack_w32_nbl10(gpu):
    """The bench workload's model at W=32, nbl=10 (polar-optical phonons,
    transverse k pair, one z bin per layer), 24 energies, 16 outer passes."""
    hip, eng, orc = gpu
    from libnegf_amd import workloads as wl
    c = wl.CI
    w = wl.config_inel(nbl=10, W=32, nE=24, nk=2)
    E = w["E"]
    shift = int(round(c["wq"] / (E[1] - E[0])))
    K = orc.inel_kmat_po(w["ndz"], 1.0, w["kfrac"], w["kw"], c["coupling"], c["wq"], c["area"], c["eps0"],
                         c["eps_inf"], c["q0"], latt=w["latt"])
    frm = np.stack([fermi(E, w["mu"][0], w["kT"]), fermi(E, w["mu"][1], w["kT"]), np.zeros_like(E)], axis=1)
    ds = hip.DeviceStructure(w["structs"])
    g = hip.solve_inelastic_mw(eng, ds, E, w["kw"], frm, 3, K, w["iz"], shift, w["Nq"], True, 15, 1e-12)
    ds.close()
    r = orc.inelastic_mw(w["structs"], w["kw"], E, 1e-4, frm, 3, K, w["iz"], shift, w["Nq"], True, 15, 1e-12)
    assert g["passes"] == r["passes"] == 16 and (g["status"] == 0).all()
    np.testing.assert_array_equal(g["iters"], r["iters_ela"])
    scale = np.abs(r["curr_mat"]).max()
    err = np.abs(g["curr_mat"] - r["curr_mat"]).max() / scale
    # Several energies sit at band edges of the stacked ribbons: there two FP64
    # evaluations of the same algorithm differ by a few 1e-10 (the oracle's
    # built-in loops against MKL: 4e-11 before the first inelastic pass, 3e-10
    # after it).  Bar: 1e-10, or four times that backend spread where MKL loads.
    bar = 1e-10
    try:
        orc.set_backend("mkl")
        rm = orc.inelastic_mw(w["structs"], w["kw"], E, 1e-4, frm, 3, K, w["iz"], shift, w["Nq"], True, 15,
                              1e-12)
        bar = max(bar, 4.0 * np.abs(rm["curr_mat"] - r["curr_mat"]).max() / scale)
    except RuntimeError:
        pass
    finally:
        orc.set_backend("builtin")
    assert bar <= 1e-8
    assert err <= bar, (err, bar)


@pytest.mark.parametrize("nk", [1, 5, 17])
def test_kpoint_counts(gpu, nk):
    """The self-energy kernel's q sums for 1 k-point, a non-power-of-two
    count held in registers, and more k-points than its register tiers (17)."""
    hip, eng, orc = gpu
    kw = np.arange(1, nk + 1, dtype=float)
    kw /= kw.sum()
    structs, E, K, iz, Nq, frm = _case(2, 2, nk, 11, 1, kw)
    _compare(hip, eng, orc, structs, kw, E, K, iz, 1, Nq, frm, True, niter=6)


def test_lanes_and_batches_bitwise(gpu):
    """Item groups on concurrent streams and smaller batches leave every
    result bitwise unchanged."""
    hip, eng, orc = gpu
    kw = np.array([0.3, 0.7])
    structs, E, K, iz, Nq, frm = _case(3, 2, 2, 25, 2, kw)
    ds = hip.DeviceStructure(structs)
    r0 = hip.solve_inelastic_mw(eng, ds, E, kw, frm, 3, K, iz, 2, Nq, True, 40, 1e-9)
    for mb, lanes in ((0, 2), (7, 2), (0, 3), (11, 1)):
        r1 = hip.solve_inelastic_mw(eng, ds, E, kw, frm, 3, K, iz, 2, Nq, True, 40, 1e-9, max_batch=mb,
                                    lanes=lanes)
        np.testing.assert_array_equal(r0["curr_mat"], r1["curr_mat"])
        np.testing.assert_array_equal(r0["iters"], r1["iters"])
        assert r0["passes"] == r1["passes"]
    ds.close()


def test_conservation_and_zero_coupling(gpu):
    hip, eng, orc = gpu
    kw = np.array([0.3, 0.7])
    structs, E, K, iz, Nq, frm = _case(3, 2, 2, 53, 2, kw, Emin=-2.6, Emax=2.6)
    ds = hip.DeviceStructure(structs)
    g = hip.solve_inelastic_mw(eng, ds, E, kw, frm, 3, K, iz, 2, Nq, True, 60, 1e-12)
    assert g["passes"] < 60
    I = g["curr_mat"]
    assert abs(I.sum()) <= 1e-8 * np.abs(I).sum()
    g0 = hip.solve_inelastic_mw(eng, ds, E, kw, frm, 3, 0.0 * K, iz, 2, Nq, True, 60, 1e-12)
    assert g0["passes"] == 2 and g0["err"] == 0.0
    ds.close()
    assert np.abs(g0["currents"] - g["currents"]).max() > 1e-3 * np.abs(g0["currents"]).max()


def test_rejects_bad_inputs(gpu):
    hip, eng, orc = gpu
    kw = np.array([0.5, 0.5])
    structs, E, K, iz, Nq, frm = _case(3, 2, 2, 9, 1, kw)
    ds = hip.DeviceStructure(structs)
    with pytest.raises(Exception, match="uniform"):
        E2 = E.copy()
        E2[3] += 0.01
        hip.solve_inelastic_mw(eng, ds, E2, kw, frm, 3, K, iz, 1, Nq, True, 3, 1e-9)
    with pytest.raises(Exception, match="K table"):
        hip.solve_inelastic_mw(eng, ds, E, kw, frm, 3, K[:2], iz, 1, Nq, True, 3, 1e-9)
    with pytest.raises(Exception, match="k-point"):
        hip.solve_inelastic_mw(eng, ds, E, np.ones(3) / 3, frm, 3, K, iz, 1, Nq, True, 3, 1e-9)
    ds.close()
