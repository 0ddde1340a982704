"""Structures whose right contact has a surface block (ns orbitals between
the device and the lead's first principal layer, surface_green,
contselfenergy.F90:186-239), and the same physical system with that surface
moved into the device as an extra last layer and a plain contact."""
import numpy as np

from libnegf_amd import workloads as wl
from libnegf_amd.structure import BlockStructure, Contact


def _herm(rng, n, scale):
    a = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    return scale * (a + a.conj().T) / 2


def _rand(rng, shape, scale):
    return scale * (rng.standard_normal(shape) + 1j * rng.standard_normal(shape))


def surface_pair(W=5, nbl=4, ns=3, seed=2, overlap=False):
    """(A, B).  A: the right contact (on block nbl) has a surface block of ns
    orbitals coupled to the device block (Hms) and to PL1 (Hs1); the device
    couples to the contact only through the surface.  B: the surface is the
    device's block nbl + 1 and the contact (plain) couples it to PL1.  The two
    have the same Green's function on the common orbitals when the device is
    evaluated at the contacts' complex energy."""
    rng = np.random.default_rng(seed)
    base = wl.ribbon(W, nbl, disorder=0.3, seed=seed)
    cR, cL = base.contacts
    Hss = _herm(rng, ns, 0.5)
    Hs1 = _rand(rng, (ns, W), 0.6)
    Hms = _rand(rng, (W, ns), 0.7)
    kwS, Sd, Su = {}, None, None
    if overlap:
        Sss = np.eye(ns) + _herm(rng, ns, 0.05)
        Ss1 = _rand(rng, (ns, W), 0.04)
        Sms = _rand(rng, (W, ns), 0.04)
        kwS = dict(Sss=Sss, Ss1=Ss1, S1s=Ss1.conj().T, SMC=np.hstack([Sms, np.zeros((W, W))]),
                   S00=np.eye(W), S01=np.zeros((W, W)))
    A_right = Contact(cblk=nbl, H00=cR.H00, H01=cR.H01, HMC=np.hstack([Hms, np.zeros((W, W))]),
                      Hss=Hss, Hs1=Hs1, H1s=Hs1.conj().T, **kwS)
    if overlap:
        cL = Contact(cblk=1, H00=cL.H00, H01=cL.H01, HMC=cL.HMC, S00=np.eye(W), S01=np.zeros((W, W)),
                     SMC=np.zeros((W, W)))
    A = BlockStructure(Hd=base.Hd, Hu=base.Hu, Hl=base.Hl, contacts=[A_right, cL], name="surfA")
    B_right = Contact(cblk=nbl + 1, H00=cR.H00, H01=cR.H01, HMC=Hs1,
                      **(dict(S00=np.eye(W), S01=np.zeros((W, W)), SMC=kwS["Ss1"]) if overlap else {}))
    if overlap:
        Sd = [np.eye(W)] * nbl + [kwS["Sss"]]
        Su = [np.zeros((W, W))] * (nbl - 1) + [Sms]
    B = BlockStructure(Hd=base.Hd + [Hss], Hu=base.Hu + [Hms], Hl=base.Hl + [Hms.conj().T],
                       contacts=[B_right, cL], Sd=Sd, Su=Su,
                       Sl=None if Su is None else [u.conj().T for u in Su], name="surfB")
    return A, B


def lead_copy_surface(W=5, nbl=4, seed=4):
    """(A, B): A's right contact has a surface block equal to one lead PL
    (Hss = H00, Hs1 = H01) with the device coupled to the surface only; B is
    the plain contact.  Both describe the same semi-infinite lead."""
    base = wl.ribbon(W, nbl, disorder=0.2, seed=seed)
    cR, cL = base.contacts
    A_right = Contact(cblk=nbl, H00=cR.H00, H01=cR.H01, HMC=np.hstack([cR.HMC, np.zeros((W, W))]),
                      Hss=cR.H00, Hs1=cR.H01, H1s=cR.H01.conj().T)
    A = BlockStructure(Hd=base.Hd, Hu=base.Hu, Hl=base.Hl, contacts=[A_right, cL], name="leadcopy")
    return A, base
