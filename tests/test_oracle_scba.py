"""CPU pins of the oracle's elastic-dephasing SCBA (or_scba_mw_sigma) and of
the decimation's SAVE'd okCo flag, by dense linear algebra that shares no code
with the block recursions.

* Fixed point.  The converged SCBA of elphdd (elphdd.F90:108-276) satisfies
  G = (E S - H - Sigma_L - Sigma_R - diag(sigma_r))^-1 with
  sigma_r = c^2 diag(G), and G^n = G (Sigma^n_contacts + diag(sigma_n)) G^dag
  with sigma_n = c^2 diag(G^n).  The second condition is LINEAR in sigma_n:
  (I - M) sigma_n = b with M_ij = c_i^2 |G_ij|^2, b = c^2 diag(G Sn_c G^dag),
  solved here directly (no iteration), so a wrong coupling scale, a dropped
  term or a transposed block in the oracle's recursion shows up at 1e-10.
* Stopping rule.  The pass-by-pass trajectory (sigma from the previous pass,
  iterative.F90:289-363) recomputed with dense inverses, and the maximum
  change of G^n over the entries in S's sparsity pattern (blk2csr,
  iterative.F90:1375-1450; check_Mat_convergence, scba.F90:120-147), give the
  pass count for a diagonal pattern and for one with off-diagonal S entries;
  the tolerance is chosen where the two differ.
* okCo.  With the flag carried from call to call (contselfenergy.F90:284) the
  counts and self-energies equal those of fresh calls whenever the first
  iterate is above SGFACC, since the flag is reset there.
"""
import numpy as np
import pytest

from libnegf_amd import workloads as wl
from libnegf_amd.structure import BlockStructure
from tests.chains import gamma


def _sigmas(oracle, s, E, delta):
    return [oracle.contact_sigma(c, s.bsz[c.cblk - 1], E + 1j * delta)[0] for c in s.contacts]


def _dense_a0(s, E, sig):
    A = s.dense_device(E)
    off = s.block_offsets()
    for c, Sg in zip(s.contacts, sig):
        b = c.cblk - 1
        A[off[b]:off[b + 1], off[b]:off[b + 1]] -= Sg
    return A, off


def _sn_contacts(s, sig, frm, ref, off):
    nd = off[-1]
    Sn = np.zeros((nd, nd), complex)
    fref = frm[ref - 1]
    for j, (c, Sg) in enumerate(zip(s.contacts, sig)):
        if j == ref - 1:
            continue
        b = c.cblk - 1
        Sn[off[b]:off[b + 1], off[b]:off[b + 1]] += (frm[j] - fref) * gamma(Sg)
    return Sn


def dense_fixed_point(s, E, sig, c2, frm, ref):
    A0, off = _dense_a0(s, E, sig)
    sr = np.zeros(off[-1], complex)
    for _ in range(2000):
        G = np.linalg.inv(A0 - np.diag(sr))
        new = c2 * np.diag(G)
        done = np.abs(new - sr).max() <= 1e-17 * max(np.abs(new).max(), 1e-300)
        sr = new
        if done:
            break
    G = np.linalg.inv(A0 - np.diag(sr))
    Snc = _sn_contacts(s, sig, frm, ref, off)
    M = c2[:, None] * np.abs(G) ** 2
    b = c2 * np.diag(G @ Snc @ G.conj().T)
    sn = np.linalg.solve(np.eye(len(b)) - M, b)
    Gn = G @ (Snc + np.diag(sn)) @ G.conj().T
    curr = []
    for j, (c, Sg) in enumerate(zip(s.contacts, sig)):
        bb = slice(off[c.cblk - 1], off[c.cblk])
        Gam = gamma(Sg)
        A = 1j * (G[bb, bb] - G[bb, bb].conj().T)
        curr.append(-np.real(np.trace(Gam @ Gn[bb, bb] - frm[j] * Gam @ A)))
    return G, Gn, np.array(curr)


def _slab_with_pattern(seed=9):
    """Small k-slab with S != I whose S blocks keep some exact zeros."""
    s = wl.KSlab.make(n=5, nbl=4, seed=seed).at_k(0.3, -0.7)
    rng = np.random.default_rng(seed)
    Sd = [a.copy() for a in s.Sd]
    for a in Sd:
        m = np.triu(rng.random(a.shape) < 0.4, 1)
        a[m | m.T] = 0.0
    return BlockStructure(Hd=s.Hd, Hu=s.Hu, Hl=s.Hl, contacts=s.contacts, Sd=Sd, Su=s.Su,
                          Sl=s.Sl, name="slab with S pattern")


@pytest.mark.parametrize("which", ["ribbon", "slab"])
def test_scba_dense_fixed_point(oracle, which):
    if which == "ribbon":
        s = wl.ribbon(5, 6, disorder=0.4, seed=23)
        Es, cpl = (-0.8, 0.45), 0.35
    else:
        s = _slab_with_pattern()
        Es, cpl = (-0.2, 0.6), 0.25
    nd = s.ndev
    cp = cpl * (1.0 + 0.5 * np.sin(np.arange(nd)))   # non-uniform coupling per orbital
    c2 = cp ** 2
    frm, ref = [0.9, 0.15, 0.0], 3
    for E in Es:
        sig = _sigmas(oracle, s, E, 1e-5)
        r = oracle.scba_mw(s, E, 1e-5, cp, 500, 1e-14, frm, ref, sigmas=sig)
        assert r["err"] < 1e-14 and r["iters"] < 500
        G, Gn, curr = dense_fixed_point(s, E, sig, c2, frm, ref)
        gr, gn = np.diag(G), np.diag(Gn)
        assert np.abs(r["gr_diag"] - gr).max() <= 1e-10 * np.abs(gr).max()
        assert np.abs(r["gn_diag"] - gn).max() <= 1e-10 * np.abs(gn).max()
        # currents against the size of the two trace terms they are the difference of
        assert np.abs(r["curr"] - curr).max() <= 1e-10 * np.abs(curr).max()
        # the coupling enters squared (elphdd.F90:94): c instead of c^2 is far off
        _, Gn_lin, _ = dense_fixed_point(s, E, sig, cp, frm, ref)
        assert np.abs(r["gn_diag"] - np.diag(Gn_lin)).max() > 1e-4 * np.abs(gn).max()


def _pattern(s):
    """Entries of G^n that blk2csr keeps (S's pattern; S = I: the diagonal),
    as a dense boolean mask."""
    off = s.block_offsets()
    nd = off[-1]
    P = np.zeros((nd, nd), bool)
    for i in range(s.nbl):
        bb = slice(off[i], off[i + 1])
        P[bb, bb] = (s.Sd[i] != 0) if s.Sd is not None else np.eye(s.bsz[i], dtype=bool)
    for i in range(s.nbl - 1):
        if s.Su is not None:
            P[off[i]:off[i + 1], off[i + 1]:off[i + 2]] = s.Su[i] != 0
        if s.Sl is not None:
            P[off[i + 1]:off[i + 2], off[i]:off[i + 1]] = s.Sl[i] != 0
    return P


def dense_trajectory(s, E, sig, c2, frm, ref, npass):
    """Per-pass maximum change of G^n over S's pattern (pass k >= 1), with the
    self-energies of pass k - 1 entering pass k."""
    A0, off = _dense_a0(s, E, sig)
    Snc = _sn_contacts(s, sig, frm, ref, off)
    P = _pattern(s)
    sr = np.zeros(off[-1], complex)
    sn = np.zeros(off[-1], complex)
    old, errs = None, []
    for k in range(npass):
        G = np.linalg.inv(A0 - np.diag(sr))
        Gn = G @ (Snc + np.diag(sn)) @ G.conj().T
        sr, sn = c2 * np.diag(G), c2 * np.diag(Gn)
        if old is not None:
            errs.append(np.abs(Gn - old)[P].max())
        old = Gn
    return np.array(errs)  # errs[k-1] = change at pass k


def _passes(errs, tol):
    for k, e in enumerate(errs, start=1):
        if e < tol:
            return k + 1
    return None


def test_scba_stopping_rule_s_pattern(oracle):
    """S's off-diagonal pattern enters the convergence test: with the same
    physics, a pattern with off-diagonal entries needs more passes than the
    diagonal one, and the oracle's pass counts match the dense trajectory."""
    base = wl.ribbon(4, 5, disorder=0.5, seed=31)
    nd = base.ndev
    # S = I exactly, and S = I plus every other entry of its tridiagonal
    # blocks at 1e-200 (their physical effect is far below rounding; their
    # pattern is not)
    tiny = 1e-200
    full = np.ones((4, 4))
    Sd = [np.eye(4) + tiny * (full - np.eye(4)) for _ in range(base.nbl)]
    Su = [tiny * full for _ in range(base.nbl - 1)]
    Sl = [tiny * full for _ in range(base.nbl - 1)]
    s_off = BlockStructure(Hd=base.Hd, Hu=base.Hu, Hl=base.Hl, contacts=base.contacts, Sd=Sd, Su=Su,
                           Sl=Sl, name="ribbon, S pattern")
    s_diag = BlockStructure(Hd=base.Hd, Hu=base.Hu, Hl=base.Hl, contacts=base.contacts,
                            Sd=[np.eye(4)] * base.nbl, Su=None, Sl=None, name="ribbon, S = I")
    cp = 0.45 * (1.0 + 0.5 * np.cos(np.arange(nd)))
    frm, ref = [1.0, 0.0, 0.0], 3
    E = 0.37
    sig = _sigmas(oracle, base, E, 1e-5)
    e_off = dense_trajectory(s_off, E, sig, cp ** 2, frm, ref, 60)
    e_diag = dense_trajectory(s_diag, E, sig, cp ** 2, frm, ref, 60)
    assert (e_off >= e_diag).all()
    # a tolerance (away from every per-pass change) at which the counts differ
    tol = None
    for k in range(len(e_diag)):
        if e_diag[k] < e_off[k] and e_diag[k] > 1e-13:
            cand = np.sqrt(e_diag[k] * e_off[k])
            gaps = np.abs(np.log(np.r_[e_off, e_diag] / cand))
            if gaps.min() > 0.02:
                tol = cand
                break
    assert tol is not None
    n_off, n_diag = _passes(e_off, tol), _passes(e_diag, tol)
    assert n_off > n_diag
    r_off = oracle.scba_mw(s_off, E, 1e-5, cp, 100, tol, frm, ref, sigmas=sig)
    r_diag = oracle.scba_mw(s_diag, E, 1e-5, cp, 100, tol, frm, ref, sigmas=sig)
    assert (r_off["iters"], r_diag["iters"]) == (n_off, n_diag)


@pytest.mark.parametrize("which", ["c1", "ribbon64", "slab48"])
def test_decimation_okco_carry_over(oracle, which):
    """contselfenergy.F90:284 keeps okCo between calls (implicit SAVE).  Along
    an energy sweep, carrying it gives the same counts and bitwise the same
    self-energies as starting each call with .false.: the first iterate of
    every call is above SGFACC, which resets the flag."""
    if which == "c1":
        s, E, eta = wl.config1()
        E = E[::17]
    elif which == "ribbon64":
        s = wl.ribbon(64, 2, disorder=0.0, seed=1)
        E, eta = np.linspace(-3.5, 3.5, 9), 1e-6
    else:
        s = wl.KSlab.make(n=48, nbl=2, seed=3).at_k(0.4, 1.1)
        E, eta = np.linspace(-2, 2, 7), 1e-5
    for c in s.contacts:
        nd = s.bsz[c.cblk - 1]
        okco = np.zeros(1, dtype=np.int32)
        for e in E:
            fresh, _, n_fresh = oracle.contact_sigma(c, nd, e + 1j * eta)
            carried, _, n_car = oracle.contact_sigma(c, nd, e + 1j * eta, okco=okco)
            assert n_car == n_fresh
            assert np.array_equal(carried, fresh)
            assert okco[0] == 1   # the flag leaves each call set
