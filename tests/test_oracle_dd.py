"""Pins of the double-double transmission (oracle/negf_dd.c), the truth the
full-size GPU accuracy tests measure against.

The same algorithm -- Sancho-Rubio decimation with the CPU stopping rule
(contselfenergy.F90:271-340), Sigma = Tlc GS Tcl (:445-491), the gsmr sweep at
real E and G11 (iterative_cpu.F90:51-198), T = |Re Tr[Gam A - Gam G Gam
G^dag]| (:399-470) -- is written here a second time in mpmath at 60 decimal
digits, with mpmath's own LU inverse.  On small ill-conditioned structures
(eta down to 1e-7, S != I) the double-double value must equal it to 1e-24
relative, or -- at band-centre points where the recursion amplifies rounding
so much that FP64 is off by 1e-2 -- be at least 1e9 times closer to it than
the FP64 oracle (T below 1e-8 is a difference of two traces of order one:
there 1e-22).  A wrong error term in a double-double primitive, a dropped low
word or a transposed
operand in the dd recursion fails here.
"""
import mpmath as mp
import numpy as np
import pytest

from libnegf_amd import workloads as wl

SGFACC = 1e-9  # ln_precision.F90:40


def _m(a):
    a = np.asarray(a)
    return mp.matrix([[mp.mpc(complex(a[i, j])) for j in range(a.shape[1])]
                      for i in range(a.shape[0])])


def _dag(A):
    return A.transpose_conj()


def _form(z, S, H, n):
    """z S - H (S None: identity)."""
    Hm = _m(H)
    Sm = _m(S) if S is not None else mp.eye(n)
    return z * Sm - Hm


def _maxabs(A):
    return max(abs(A[i, j]) for i in range(A.rows) for j in range(A.cols))


def _sigma_mp(c, n, E, delta):
    z = mp.mpc(E, delta)
    zc = mp.conj(z)
    S00 = c.S00
    S01 = c.S01 if c.S01 is not None else np.zeros_like(c.H01)
    SMC = c.SMC if c.SMC is not None else np.zeros_like(c.HMC)
    Ao = _form(z, S00, c.H00, n)
    Bo = z * _m(S01) - _m(c.H01)
    Co = _dag(zc * _m(S01) - _m(c.H01))
    Aos = Ao.copy()
    okco, ncyc = False, 0
    for it in range(1, 301):
        ncyc = it
        Go = mp.inverse(Ao)
        GoXCo = Go * Co
        C1 = Co * GoXCo
        if _maxabs(C1) <= SGFACC:
            if okco:
                break
            okco = True
        else:
            okco = False
        Self = Bo * GoXCo
        Aos -= Self
        Ao -= Self
        GoXBo = Go * Bo
        B1 = Bo * GoXBo
        Ao -= Co * GoXBo
        Bo, Co = B1, C1
    else:
        ncyc = 301
    GS = mp.inverse(Aos)
    Tlc = z * _m(SMC) - _m(c.HMC)
    Tcl = _dag(zc * _m(SMC) - _m(c.HMC))
    return Tlc * GS * Tcl, ncyc


def transmission_mp(s, E, delta):
    n, nbl = s.bsz[0], s.nbl
    sig, ncyc = [], []
    for c in s.contacts:
        Sg, k = _sigma_mp(c, n, E, delta)
        sig.append(Sg)
        ncyc.append(k)
    E = mp.mpf(E)

    def esh_diag(i):
        D = _form(E, s.Sd[i] if s.Sd is not None else None, s.Hd[i], n)
        for c, Sg in zip(s.contacts, sig):
            if c.cblk == i + 1:
                D -= Sg
        return D

    def off(S, H, i):
        return E * _m(S[i]) - _m(H[i]) if S is not None else -_m(H[i])

    g = mp.inverse(esh_diag(nbl - 1))
    for i in range(nbl - 2, -1, -1):
        g = mp.inverse(esh_diag(i) - off(s.Su, s.Hu, i) * g * off(s.Sl, s.Hl, i))
    ct1 = [c.cblk for c in s.contacts].index(1)
    Sg = sig[ct1]
    Gam = mp.mpc(0, 1) * (Sg - _dag(Sg))
    A = mp.mpc(0, 1) * (g - _dag(g))
    X = Gam * A - Gam * g * Gam * _dag(g)
    return abs(sum(X[i, i] for i in range(n)).real), ncyc


CASES = [
    ("slab S!=I", lambda: wl.KSlab.make(n=3, nbl=4, seed=11).at_k(0.4, -0.9), 1e-6),
    ("slab S=I", lambda: wl.KSlab.make(n=4, nbl=3, seed=5, with_overlap=False).at_k(1.1, 0.2), 1e-7),
    ("disordered ribbon", lambda: wl.ribbon(3, 5, disorder=0.8, seed=2), 1e-6),
]


@pytest.mark.parametrize("name,make,eta", CASES, ids=[c[0] for c in CASES])
def test_dd_against_mpmath(oracle, name, make, eta):
    mp.mp.dps = 60
    s = make()
    E = np.linspace(-2.3, 2.3, 9)
    th, tl, ncyc = oracle.transmission_dd(s, E, eta)
    T64, n64, _ = oracle.transmission_batch(s, E, eta)
    worst_dd, worst_64 = 0.0, 0.0
    for i, e in enumerate(E):
        Tm, nm = transmission_mp(s, float(e), eta)
        assert list(ncyc[i]) == nm and list(n64[i]) == nm
        if Tm < 1e-16:
            continue
        d = float(abs((mp.mpf(th[i]) + mp.mpf(tl[i]) - Tm) / Tm))
        e64 = float(abs((mp.mpf(T64[i]) - Tm) / Tm))
        assert d <= max(1e-24 if Tm >= 1e-8 else 1e-22, 1e-9 * e64), (name, e, d, e64)
        worst_dd = max(worst_dd, d)
        worst_64 = max(worst_64, e64)
    # ill-conditioned enough that FP64 rounding shows
    assert worst_64 > 1e-12, (name, worst_64)
