"""The oracle's electron/hole density of compute_density_efa
(libnegf.F90:1690-1791): contour_int_n_def / contour_int_p_def
(integrations.F90:281-592), trapez (:1529-1547), real_axis_int_{n,p}_def
(:1346-1462) and the hole integrand A - G^n (compute_Gp, :2405-2420), pinned
against:

* a single level: the contour sum of w / (z - e) gives g f(e) for electrons
  and g (1 - f(e)) for holes;
* the trapezoid rule's exactness on linear functions;
* the ideal chain at finite temperature: n_i = 2 int rho(E) f(E) dE and
  p_i = 2 int rho(E) (1 - f(E)) dE (rho = 1 / (pi sqrt(4 - E^2))), so
  n_i + p_i = 2;
* the hole real-axis term: with every contact level below the window the
  occupations vanish and p_i is the full band (2), with every level above it
  p_i vanishes.
"""
import numpy as np
import pytest
from scipy.integrate import quad

from tests.chains import chain_structure

P0 = dict(mu=[0.0, 0.0], mu_n=[0.0, 0.0], mu_p=[0.0, 0.0], min_or_max=1, np_n=[60, 150], np_p=[60, 150],
          n_poles=4, ec=-3.0, deltaec=0.0, ev=3.0, deltaev=0.0, kbt_dm=[0.05, 0.05], n_kt=30,
          g_spin=2.0, np_real=0, delta=1e-4)


def _fermi(E, mu, kT):
    return 1.0 / (np.exp((E - mu) / kT) + 1.0)


@pytest.mark.parametrize("particle", [1, -1])
@pytest.mark.parametrize("e", [-1.3, -0.02, 0.05, 0.9])
def test_single_level_contours(oracle, particle, e):
    mu, kT = 0.1, 0.05
    Eb = -2.0 if particle == 1 else 2.0
    z, w = oracle.efa_contour_grid(particle, Eb, mu, kT, 30, 60, 150, 4, 2.0)
    assert len(z) == 214
    T = np.sum(w / (z - e))
    occ = 1j * (T - np.conj(T))  # zspectral
    f = _fermi(e, mu, kT)
    want = 2.0 * (f if particle == 1 else 1.0 - f)
    assert abs(occ.real - want) < 1e-11 and abs(occ.imag) < 1e-12


def test_contour_edge_cases(oracle):
    # an edge within 1e-3 of muref moves to muref -+ kT; an edge beyond muref drops the poles
    z, _ = oracle.efa_contour_grid(1, 0.0005, 0.0, 0.01, 10, 4, 4, 2, 2.0)
    assert abs(z[0].real + 0.01) < 1e-15 and len(z) == 10
    z, _ = oracle.efa_contour_grid(1, 0.5, 0.0, 0.01, 10, 4, 4, 2, 2.0)
    assert len(z) == 8
    z, _ = oracle.efa_contour_grid(-1, -0.5, 0.0, 0.01, 10, 4, 4, 2, 2.0)
    assert len(z) == 8


def test_trapez(oracle):
    x, w = oracle.trapez(-1.5, 2.5, 9)
    assert abs(w.sum() - 4.0) < 1e-15 and abs(np.dot(w, 3 * x - 1) - (3 * (2.5**2 - 1.5**2) / 2 - 4)) < 1e-13
    assert x[0] == -1.5 and x[-1] == 2.5 and abs(w[0] - 0.25) < 1e-16


def _efa(oracle, particle, **kw):
    s = chain_structure()
    N = s.ndev + 40
    p = dict(P0, **kw)
    return oracle.density_efa(s, N, [60, 80], np.eye(N, dtype=bool), particle, p)


def test_chain_electrons_holes_closed_form(oracle):
    mu, kT = 0.3, 0.05
    n_ref = 2 * quad(lambda t: _fermi(2 * np.cos(t), mu, kT) / np.pi, 0, np.pi, limit=500, epsabs=1e-13)[0]
    kw = dict(mu=[mu, mu], mu_n=[mu, mu], mu_p=[mu, mu], kbt_dm=[kT, kT])
    _, _, n = _efa(oracle, 1, **kw)
    _, _, p = _efa(oracle, -1, **kw)
    assert np.abs(n[:60] - n_ref).max() < 1e-8
    assert np.abs(p[:60] - (2.0 - n_ref)).max() < 1e-8
    assert np.abs(n[:60] + p[:60] - 2.0).max() < 1e-8


def test_chain_hole_real_axis_limits(oracle):
    # no contour: real axis only, reference contact ncont + 1 (occupation 0)
    kw = dict(np_p=[0, 0], n_poles=0, np_real=4001, delta=2e-3, mu_p=[-2.6, -2.6], ev=2.6, deltaev=0.0,
              n_kt=0, kbt_dm=[1e-3, 1e-3])
    _, _, p_low = _efa(oracle, -1, mu=[-2.7, -2.7], **kw)   # contacts empty: p = band weight
    _, _, p_high = _efa(oracle, -1, mu=[2.7, 2.7], **kw)    # contacts full: p = O(delta) weight
    assert np.abs(p_low[:60] - 2.0).max() < 2e-2
    assert p_high[:60].max() < 2e-2 and p_high[:60].min() > -1e-10


# ---- quasi-equilibrium densities (quasiEq_int_n / _p) ----
def _qp(**kw):
    p = dict(P0, np_n=[60, 300], np_p=[60, 300], kwght=1.0)
    p.update(kw)
    return p


def test_quasi_eq_chain_closed_form(oracle):
    """Uniform levels on the ideal chain: the windows cover the band, so
    n = 2 int rho f and p = 2 - n per site."""
    s = chain_structure()
    mu, kT = 0.3, 0.05
    n_ref = 2 * quad(lambda t: _fermi(2 * np.cos(t), mu, kT) / np.pi, 0, np.pi, limit=500, epsabs=1e-13)[0]
    nz = 60
    one = np.ones(nz)
    p = _qp(kbt_dm=[kT, kT], deltaec=5.2, deltaev=5.2)
    n = oracle.density_quasi(s, 1, p, 2.2 * one, -2.2 * one, mu * one, mu * one)
    h = oracle.density_quasi(s, -1, p, 2.2 * one, -2.2 * one, mu * one, mu * one)
    assert np.abs(n - n_ref).max() < 1e-8
    assert np.abs(h - (2.0 - n_ref)).max() < 1e-8


def test_quasi_eq_per_orbital_levels(oracle):
    """An orbital's density depends on its own level only: a two-level
    profile equals, site by site, the uniform runs at each level."""
    s = chain_structure()
    nz = 60
    one = np.ones(nz)
    p = _qp(kbt_dm=[0.05, 0.05], deltaec=5.2, deltaev=5.2, np_n=[20, 60], np_p=[20, 60])
    lev = np.where(np.arange(nz) < 25, -0.4, 0.6)
    for part in (1, -1):
        prof = oracle.density_quasi(s, part, p, 2.2 * one, -2.2 * one, lev, lev)
        for m in (-0.4, 0.6):
            uni = oracle.density_quasi(s, part, p, 2.2 * one, -2.2 * one, m * one, m * one)
            sel = lev == m
            assert np.abs(prof[sel] - uni[sel]).max() < 1e-13
