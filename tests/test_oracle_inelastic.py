"""The oracle's inelastic electron-phonon SCBA (elphinel.F90, self_energy.F90,
meir_wingreen in integrations.F90:1804-1966), pinned by:

  * the polar-optical K table against a numerical q_z quadrature of the
    screened Froehlich coupling (q^2 / (q^2 + q0^2)^2 with q^2 = Q^2 + q_z^2),
    whose closed form prepare_POKmat tabulates;
  * the non-polar K table: its z profile against the quadrature of a
    Lorentzian (exp(-|q| z)), its z = 0 value against D0 cos^2(k, q) / 2 with
    the absolute k-points from numpy's inverse of the lattice;
  * a dense brute-force restatement of the outer SCBA loop (whole-device
    inverses, Sigma assembled on a dense band mask) -- same currents, same
    pass count, with and without the off-diagonal self-energy blocks;
  * particle conservation of the scattering self-energies: for ANY G^r / G^n
    blocks, sum over (E, k) of w_k Tr[Sigma_n A - Gamma_s G^n] vanishes when
    Sigma_n and Sigma_r come from the same G (Gamma_s = i (Sigma_r - Sigma_r^H));
    it fails when the two use inconsistent phonon factors;
  * at convergence the contact currents summed over energy and contacts vanish;
  * with zero coupling the loop stops after two passes with the ballistic current.
"""
import numpy as np
import pytest
from scipy import integrate

from oracle import pyoracle
from tests.inel_model import LATT, band_mask, bose, dense_inelastic, fermi, k_system


@pytest.fixture(scope="module")
def orc():
    pyoracle.build()
    pyoracle.set_backend("builtin")
    return pyoracle


def _kabs(kfrac):
    rec = 2 * np.pi * np.linalg.inv(LATT).T   # reciprocal vectors as columns
    return kfrac @ rec.T


def test_kmat_po_quadrature(orc):
    kfrac = np.array([[0.0, 0.0, 0.0], [0.2, 0.1, 0.0], [0.45, 0.3, 0.0]])
    kw = np.array([0.2, 0.5, 0.3])
    q0, dz, ndz = 0.7, 0.35, 9
    coupling, wq, area, eps0, eps_inf = 1.3, 0.06, 2.5, 12.0, 10.0
    K = orc.inel_kmat_po(ndz, dz, kfrac, kw, coupling, wq, area, eps0, eps_inf, q0, latt=LATT)
    Ce = coupling * wq / 2.0 / area * (1.0 / eps_inf - 1.0 / eps0)
    ka = _kabs(kfrac)
    for iK in range(3):
        for iQ in range(3):
            Q2 = float(np.sum((ka[iK] - ka[iQ]) ** 2))
            b2 = Q2 + q0 * q0
            for iz in (0, 1, 4, 8):
                z = iz * dz
                f = lambda qz: (Q2 + qz * qz) / (qz * qz + b2) ** 2
                if z == 0.0:
                    val = integrate.quad(f, 0, np.inf, epsabs=1e-14, epsrel=1e-12)[0] / np.pi
                else:
                    val = integrate.quad(f, 0, np.inf, weight="cos", wvar=z)[0] / np.pi
                ref = Ce * kw[iQ] * val
                # bar relative to the z = 0 value: the two terms nearly cancel where 1 - b z ~ 0
                assert abs(K[iz, iK, iQ] - ref) <= 1e-9 * K[0, iK, iQ], (iK, iQ, iz, K[iz, iK, iQ], ref)


def test_kmat_nonpo_profile(orc):
    kfrac = np.array([[0.0, 0.0, 0.0], [0.2, 0.1, 0.0], [0.45, -0.3, 0.0]])
    kw = np.array([0.25, 0.35, 0.4])
    D0, coupling, dz, ndz = 2.0, 0.8, 0.5, 6
    K = orc.inel_kmat_nonpo(ndz, dz, kfrac, kw, coupling, D0, latt=LATT)
    ka = _kabs(kfrac)
    assert np.all(K[:, 0, :] == 0.0) and np.all(K[:, :, 0] == 0.0)   # k = 0 or q = 0
    for iK in (1, 2):
        for iQ in (1, 2):
            kk, kq = ka[iK], ka[iQ]
            cos2 = np.dot(kk, kq) ** 2 / (np.dot(kk, kk) * np.dot(kq, kq))
            assert abs(K[0, iK, iQ] - coupling * kw[iQ] * D0 * cos2 / 2) <= 1e-13 * K[0, iK, iQ]
            a = np.linalg.norm(kq)
            for iz in (1, 3, 5):
                prof = integrate.quad(lambda q: 2 * a / (q * q + a * a), 0, np.inf, weight="cos",
                                      wvar=iz * dz)[0] / np.pi
                assert abs(K[iz, iK, iQ] / K[0, iK, iQ] - prof) <= 1e-8


def _setup(nk=2, nbl=3, n=2, ne=25, shift=2, coupling=3.0, tridiag=True, kw=None):
    structs, kfrac = k_system(nbl=nbl, n=n, nk=nk)
    kw = np.full(nk, 1.0 / nk) if kw is None else np.asarray(kw, float)
    E = np.linspace(-1.5, 1.5, ne)
    wq = shift * (E[1] - E[0])
    ndev = nbl * n
    iz = np.arange(ndev)            # one z bin per site along the chain, dz = 1
    K = pyoracle.inel_kmat_po(ndev, 1.0, kfrac, kw, coupling, wq, 1.0, 2.0, 1.0, 0.5, latt=LATT)
    Nq = bose(wq, 0.2)
    frm = np.stack([fermi(E, 0.3, 0.05), fermi(E, -0.3, 0.05), np.zeros_like(E)], axis=1)
    return structs, kw, E, K, iz, shift, Nq, frm


@pytest.mark.parametrize("tridiag", [True, False])
def test_inelastic_matches_dense_bruteforce(orc, tridiag):
    structs, kw, E, K, iz, shift, Nq, frm = _setup(tridiag=tridiag, kw=[0.3, 0.7])
    delta, tol, niter = 1e-4, 1e-9, 30
    r = orc.inelastic_mw(structs, kw, E, delta, frm, 3, K, iz, shift, Nq, tridiag, niter, tol)
    sigma_of = lambda s, c, z: orc.contact_sigma(s.contacts[c], s.bsz[s.contacts[c].cblk - 1], z)[0]
    cm, J, p = dense_inelastic(structs, kw, E, delta, frm, 3, K, iz, shift, Nq, tridiag, niter, tol,
                               sigma_of)
    assert r["passes"] == p
    assert 4 <= p < niter                       # converged, after several passes
    scale = np.abs(cm).max()
    assert np.abs(r["curr_mat"] - cm).max() <= 1e-10 * scale
    assert np.abs(r["currents"] - J).max() <= 1e-10 * np.abs(J).max()


def _trace_blocks(bsz, Sd, Su, Sl, Xd, Xu, Xl, tridiag):
    """Tr[Sigma X] of block-tridiagonal Sigma (d, u, l lists) and X."""
    t = sum(np.trace(Sd[i] @ Xd[i]) for i in range(len(bsz)))
    if tridiag:
        t += sum(np.trace(Su[i] @ Xl[i]) + np.trace(Sl[i] @ Xu[i]) for i in range(len(bsz) - 1))
    return t


def _blocks(flat, bsz, p, kind):
    nd = sum(b * b for b in bsz)
    no = sum(bsz[i] * bsz[i + 1] for i in range(len(bsz) - 1))
    out, o = [], 0
    if kind == "d":
        base = flat[p * nd:(p + 1) * nd]
        for b in bsz:
            out.append(base[o:o + b * b].reshape((b, b), order="F"))
            o += b * b
    else:
        base = flat[p * no:(p + 1) * no]
        for i in range(len(bsz) - 1):
            r, c = (bsz[i], bsz[i + 1]) if kind == "u" else (bsz[i + 1], bsz[i])
            out.append(base[o:o + r * c].reshape((r, c), order="F"))
            o += r * c
    return out


def _scattering_identity(st, bsz, kw, ne, nk, tridiag, Srd=None, Sru=None, Srl=None):
    H = lambda m: m.conj().T
    tot, mag = 0.0, 0.0
    for k in range(nk):
        for e in range(ne):
            p = k * ne + e
            g = {n: _blocks(st[n], bsz, p, n[-1]) for n in ("Grd", "Gru", "Grl", "Gnd", "Gnu", "Gnl")}
            srd = _blocks(st["Srd"] if Srd is None else Srd, bsz, p, "d")
            sru = _blocks(st["Sru"] if Sru is None else Sru, bsz, p, "u")
            srl = _blocks(st["Srl"] if Srl is None else Srl, bsz, p, "l")
            snd = _blocks(st["Snd"], bsz, p, "d")
            snu = _blocks(st["Snu"], bsz, p, "u")
            snl = [H(x) for x in snu]
            Ad = [1j * (x - H(x)) for x in g["Grd"]]
            Au = [1j * (u - H(l)) for u, l in zip(g["Gru"], g["Grl"])]
            Al = [1j * (l - H(u)) for u, l in zip(g["Gru"], g["Grl"])]
            Gd = [1j * (x - H(x)) for x in srd]
            Gu = [1j * (u - H(l)) for u, l in zip(sru, srl)]
            Gl = [1j * (l - H(u)) for u, l in zip(sru, srl)]
            t1 = _trace_blocks(bsz, snd, snu, snl, Ad, Au, Al, tridiag)
            t2 = _trace_blocks(bsz, Gd, Gu, Gl, g["Gnd"], g["Gnu"], g["Gnl"], tridiag)
            tot += kw[k] * (t1 - t2)
            mag += kw[k] * (abs(t1) + abs(t2))
    return tot, mag


@pytest.mark.parametrize("tridiag", [True, False])
def test_scattering_selfenergies_conserve_particles(orc, tridiag):
    structs, kw, E, K, iz, shift, Nq, frm = _setup(tridiag=tridiag, kw=[0.3, 0.7])
    r = orc.inelastic_mw(structs, kw, E, 1e-4, frm, 3, K, iz, shift, Nq, tridiag, 2, 1e-14, stores=True)
    st, bsz = r["stores"], list(structs[0].bsz)
    tot, mag = _scattering_identity(st, bsz, kw, len(E), len(structs), tridiag)
    assert mag > 1e-3
    assert abs(tot) <= 1e-13 * mag, (tot, mag)
    # Sigma_r built with phonon factors inconsistent with Sigma_n breaks it
    ne, nk = len(E), len(structs)
    g = [st[n] for n in ("Grd", "Gru", "Grl", "Gnd", "Gnu", "Gnl")]
    bad = orc.inel_sigma(bsz, ne, nk, K, iz, shift, Nq + 0.25, tridiag, g)
    good = orc.inel_sigma(bsz, ne, nk, K, iz, shift, Nq, tridiag, g)
    for a, n in zip(good, ("Srd", "Sru", "Srl", "Snd", "Snu")):   # the driver's own Sigma
        np.testing.assert_array_equal(a, st[n])
    tot2, mag2 = _scattering_identity(st, bsz, kw, ne, nk, tridiag, Srd=bad[0], Sru=bad[1], Srl=bad[2])
    assert abs(tot2) > 1e-4 * mag2


def test_converged_currents_conserve(orc):
    structs, kw, E, K, iz, shift, Nq, frm = _setup(tridiag=True, kw=[0.3, 0.7])
    # a wider grid so that G is small at its ends
    E = np.linspace(-2.6, 2.6, 53)
    frm = np.stack([fermi(E, 0.3, 0.05), fermi(E, -0.3, 0.05), np.zeros_like(E)], axis=1)
    r = orc.inelastic_mw(structs, kw, E, 1e-4, frm, 3, K, iz, shift, Nq, True, 60, 1e-12)
    assert r["passes"] < 60
    I = r["curr_mat"]
    tot = I.sum()
    assert abs(tot) <= 1e-8 * np.abs(I).sum(), (tot, np.abs(I).sum())
    # the coupling matters: it moves the current well past that bar
    r0 = orc.inelastic_mw(structs, kw, E, 1e-4, frm, 3, 0.0 * K, iz, shift, Nq, True, 60, 1e-12)
    assert np.abs(r0["currents"] - r["currents"]).max() > 1e-3 * np.abs(r0["currents"]).max()


def test_zero_coupling_is_ballistic(orc):
    structs, kw, E, K, iz, shift, Nq, frm = _setup()
    r = orc.inelastic_mw(structs, kw, E, 1e-4, frm, 3, 0.0 * K, iz, shift, Nq, True, 10, 1e-12)
    assert r["passes"] == 2 and r["err"] == 0.0
    acc = np.zeros_like(r["curr_mat"])
    for k, s in enumerate(structs):
        for e, En in enumerate(E):
            d = orc.scba_mw(s, En, 1e-4, np.zeros(s.ndev), 0, 1e-7, frm[e], 3)
            acc[e] += kw[k] * d["curr"]
    np.testing.assert_array_equal(r["curr_mat"], acc)
