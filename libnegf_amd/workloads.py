"""Seeded synthetic inputs for the five reference workloads.

Every generator is deterministic (numpy PCG64 with the stated seed) so the
oracle, the tests and the benchmark see identical matrices.

* ``ribbon``      -- nearest-neighbour square-lattice ribbon, W orbitals per
                     principal layer (PL), hopping t, optional seeded on-site
                     disorder U[-w, w] in the device PLs only, clean contacts
                     identical to the device PL, S = I.
                     C1: W=10,  nbl=20,  no disorder, E in [-3,3] (200), eta 1e-5
                     C2: W=64,  nbl=100, w=0.5,        E in [-3.5,3.5] (2000), eta 1e-6
                     C4: W=128, nbl=30,  w=0.2 (SCBA), 1000 E, eta 1e-5
* ``KSlab``       -- C3/C5: n orbitals per PL, real-symmetric Gaussian blocks
                     with transverse Bloch phases on a kgrid x kgrid
                     Monkhorst-Pack grid.  Per-PL on-site A0_i (sigma 0.3,
                     symmetrised), per-interface coupling A1_i (sigma 0.1),
                     shared transverse couplings Bx, By (sigma 0.1);
                     H_ii(k) = A0_i + Bx e^{ikx} + Bx^T e^{-ikx} + By e^{iky} + By^T e^{-iky}.
                     S(k) = I + 0.05 R(k), R built the same way from its own
                     seeded Gaussians, scaled so 0.05*||R(k)||_2 <= 0.5 for
                     every k (S positive definite).  Contacts repeat one fixed
                     (A0c, A1c, R0c, R1c) pair and couple to the device with A1c.
                     eta 1e-5, 500 E in [-2, 2].

Contact ordering follows the reference tests: contact 1 attaches to the last
block (cblk = nbl), contact 2 to block 1 (``tests/f90_transmission/hello.F90:48``).
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import List, Optional, Tuple

import numpy as np

from .structure import BlockStructure, Contact

# ---- configuration constants ------------------------------------------------
C1 = dict(W=10, nbl=20, disorder=0.0, seed=1, nE=200, Emin=-3.0, Emax=3.0, eta=1e-5)
C2 = dict(W=64, nbl=100, disorder=0.5, seed=2, nE=2000, Emin=-3.5, Emax=3.5, eta=1e-6)
C3 = dict(n=256, nbl=40, kgrid=16, seed=3, nE=500, Emin=-2.0, Emax=2.0, eta=1e-5)
KB_EV = 8.617332411853536e-5
C4 = dict(W=128, nbl=30, disorder=0.2, seed=4, nE=1000, eta=1e-5, coupling=0.1,
          bias=0.2, T=300.0, scba_tol=1e-8, scba_niter=100)


# Inelastic (polar-optical phonon) SCBA workload: a stack of disordered
# ribbons coupled transversally with hopping ty (Bloch phase 2 ty cos(2 pi ky)
# on every site at nk transverse k-points), optical phonon wq, Froehlich
# screening q0, one z bin per principal layer.
CI = dict(W=64, nbl=20, disorder=0.2, seed=6, ty=-0.3, nk=4, nE=400, eta=1e-5, bias=0.2, T=300.0,
          wq=0.06, Tph=300.0, coupling=40.0, eps0=12.9, eps_inf=10.9, q0=0.1, area=1.0,
          tol=1e-6, niter=60)


def energy_grid(nE: int, Emin: float, Emax: float) -> np.ndarray:
    """E_i = Emin + (i-1) (Emax-Emin)/(nE-1), i = 1..nE (inclusive grid)."""
    return Emin + (Emax - Emin) * np.arange(nE, dtype=np.float64) / (nE - 1)


def _ribbon_pl(W: int, t: float) -> np.ndarray:
    H = np.zeros((W, W))
    for i in range(W - 1):
        H[i, i + 1] = H[i + 1, i] = t
    return H


def ribbon(W: int, nbl: int, disorder: float = 0.0, seed: int = 0, t: float = -1.0,
           name: str = "ribbon") -> BlockStructure:
    rng = np.random.Generator(np.random.PCG64(seed))
    H00 = _ribbon_pl(W, t)
    T = t * np.eye(W)
    Hd = []
    for _ in range(nbl):
        onsite = rng.uniform(-disorder, disorder, W) if disorder > 0 else np.zeros(W)
        Hd.append(H00 + np.diag(onsite))
    Hu = [T.copy() for _ in range(nbl - 1)]
    Hl = [T.T.copy() for _ in range(nbl - 1)]
    c1 = Contact(cblk=nbl, H00=H00, H01=T, HMC=T)        # right contact
    c2 = Contact(cblk=1, H00=H00, H01=T.T, HMC=T.T)      # left contact
    return BlockStructure(Hd=Hd, Hu=Hu, Hl=Hl, contacts=[c1, c2], name=name)


def config1() -> Tuple[BlockStructure, np.ndarray, float]:
    c = C1
    return (ribbon(c["W"], c["nbl"], c["disorder"], c["seed"], name="C1"),
            energy_grid(c["nE"], c["Emin"], c["Emax"]), c["eta"])


def config2(nbl: Optional[int] = None, W: Optional[int] = None):
    c = C2
    return (ribbon(W or c["W"], nbl or c["nbl"], c["disorder"], c["seed"], name="C2"),
            energy_grid(c["nE"], c["Emin"], c["Emax"]), c["eta"])


def config_inel(nbl: Optional[int] = None, W: Optional[int] = None, nE: Optional[int] = None,
                nk: Optional[int] = None):
    """Structures per transverse k, uniform energy grid, k-points (fractional,
    along y) and weights, lattice vectors (columns), z bins per device orbital,
    the phonon's energy shift in grid steps and Bose occupation, contact
    chemical potentials and kT."""
    c = CI
    W, nbl, nE, nk = W or c["W"], nbl or c["nbl"], nE or c["nE"], nk or c["nk"]
    base = ribbon(W, nbl, c["disorder"], c["seed"], name="CI")
    ky = (np.arange(nk) + 0.5) / (2 * nk)  # half zone (time-reversal pairs)
    structs = []
    for j in range(nk):
        d = 2.0 * c["ty"] * np.cos(2 * np.pi * ky[j]) * np.eye(W)
        cs = [Contact(cblk=ct.cblk, H00=ct.H00 + d, H01=ct.H01, HMC=ct.HMC) for ct in base.contacts]
        structs.append(BlockStructure(Hd=[h + d for h in base.Hd], Hu=base.Hu, Hl=base.Hl, contacts=cs,
                                      name="CI"))
    kfrac = np.stack([np.zeros(nk), ky, np.zeros(nk)], axis=1)
    kw = np.full(nk, 1.0 / nk)
    latt = np.diag([1.0, 1.0, 1.0])
    kT = KB_EV * c["T"]
    mu = (c["bias"] / 2.0, -c["bias"] / 2.0)
    E = energy_grid(nE, -c["bias"] / 2 - 10 * kT, c["bias"] / 2 + 10 * kT)
    shift = int(round(c["wq"] / (E[1] - E[0])))
    kph = KB_EV * c["Tph"]
    Nq = 1.0 / (np.exp(c["wq"] / kph) - 1.0)
    iz = np.repeat(np.arange(nbl), W)
    return dict(structs=structs, E=E, kfrac=kfrac, kw=kw, latt=latt, iz=iz, ndz=nbl, shift=shift,
                Nq=Nq, mu=mu, kT=kT, eta=c["eta"])


def config4(nbl: Optional[int] = None, W: Optional[int] = None):
    c = C4
    s = ribbon(W or c["W"], nbl or c["nbl"], c["disorder"], c["seed"], name="C4")
    kT = KB_EV * c["T"]
    mu = (c["bias"] / 2.0, -c["bias"] / 2.0)
    Emin, Emax = -c["bias"] / 2 - 10 * kT, c["bias"] / 2 + 10 * kT
    return s, energy_grid(c["nE"], Emin, Emax), c["eta"], mu, kT


def analytic_ribbon_T(W: int, E: np.ndarray, t: float = -1.0) -> np.ndarray:
    """Clean ribbon: T(E) = #{m : |E - eps_m| < 2|t|}, eps_m = 2 t cos(m pi/(W+1))."""
    eps = 2 * t * np.cos(np.arange(1, W + 1) * np.pi / (W + 1))
    return np.array([np.sum(np.abs(e - eps) < 2 * abs(t)) for e in np.atleast_1d(E)], float)


def analytic_chain_surface_g(z: complex, eps: float, t: float = -1.0) -> complex:
    """Surface GF of a semi-infinite 1D chain with on-site eps, hopping t:
    g = [(z-eps) - sqrt((z-eps)^2 - 4t^2)] / (2 t^2) on the branch |t g| < 1."""
    w = z - eps
    r = np.sqrt(w * w - 4 * t * t + 0j)
    g1 = (w - r) / (2 * t * t)
    g2 = (w + r) / (2 * t * t)
    return g1 if abs(t * g1) < 1 else g2


# ---- k-resolved slab (C3, C5) ---------------------------------------------------
def mp_grid(kgrid: int) -> np.ndarray:
    """Monkhorst-Pack points k_j = 2 pi ((j + 1/2)/kgrid - 1/2), as (kx, ky) pairs."""
    k1 = 2 * np.pi * ((np.arange(kgrid) + 0.5) / kgrid - 0.5)
    kx, ky = np.meshgrid(k1, k1, indexing="ij")
    return np.stack([kx.ravel(), ky.ravel()], axis=1)


def _sym_gauss(rng, n, sigma):
    a = rng.normal(0.0, sigma, (n, n))
    return (a + a.T) / math.sqrt(2.0)


@dataclass
class KSlab:
    """Real-space description of the k-resolved slab (C3/C5).  ``at_k`` builds
    the Bloch-summed BlockStructure for one transverse k = (kx, ky)."""
    n: int
    nbl: int
    A0: List[np.ndarray]
    A1: List[np.ndarray]
    Bx: np.ndarray
    By: np.ndarray
    R0: List[np.ndarray]
    R1: List[np.ndarray]
    Rx: np.ndarray
    Ry: np.ndarray
    A0c: np.ndarray
    A1c: np.ndarray
    R0c: np.ndarray
    R1c: np.ndarray
    s_scale: float
    with_overlap: bool = True

    @staticmethod
    def make(n: int = 256, nbl: int = 40, seed: int = 3, with_overlap: bool = True) -> "KSlab":
        rng = np.random.Generator(np.random.PCG64(seed))
        A0 = [_sym_gauss(rng, n, 0.3) for _ in range(nbl)]
        A1 = [rng.normal(0.0, 0.1, (n, n)) for _ in range(nbl - 1)]
        Bx = rng.normal(0.0, 0.1, (n, n))
        By = rng.normal(0.0, 0.1, (n, n))
        A0c = _sym_gauss(rng, n, 0.3)
        A1c = rng.normal(0.0, 0.1, (n, n))
        R0 = [_sym_gauss(rng, n, 1.0) for _ in range(nbl)]
        R1 = [rng.normal(0.0, 1.0, (n, n)) for _ in range(nbl - 1)]
        Rx = rng.normal(0.0, 1.0, (n, n))
        Ry = rng.normal(0.0, 1.0, (n, n))
        R0c = _sym_gauss(rng, n, 1.0)
        R1c = rng.normal(0.0, 1.0, (n, n))
        # bound on ||R(k)||_2 over all k for the full block-tridiagonal R
        nrm = lambda a: np.linalg.norm(a, 2)
        bloch = 2 * nrm(Rx) + 2 * nrm(Ry)
        on = max(max(nrm(a) for a in R0 + [R0c]) + bloch, 0.0)
        off = max(nrm(a) for a in R1 + [R1c])
        bound = on + 2 * off
        s_scale = 0.05 * min(1.0, 10.0 / bound)   # 0.05 * (scaled R) has norm <= 0.5
        return KSlab(n, nbl, A0, A1, Bx, By, R0, R1, Rx, Ry, A0c, A1c, R0c, R1c, s_scale,
                     with_overlap)

    def _onsite(self, base, bx, by, kx, ky):
        return (base + bx * np.exp(1j * kx) + bx.T * np.exp(-1j * kx)
                + by * np.exp(1j * ky) + by.T * np.exp(-1j * ky))

    def at_k(self, kx: float, ky: float, nbl: Optional[int] = None) -> BlockStructure:
        nbl = nbl or self.nbl
        s = self.s_scale
        n = self.n
        Hd = [self._onsite(self.A0[i], self.Bx, self.By, kx, ky) for i in range(nbl)]
        Hu = [self.A1[i] for i in range(nbl - 1)]
        Hl = [self.A1[i].T for i in range(nbl - 1)]
        H00 = self._onsite(self.A0c, self.Bx, self.By, kx, ky)
        Sd = Su = Sl = None
        S00 = S01r = S01l = SMCr = SMCl = None
        if self.with_overlap:
            I = np.eye(n)
            Sd = [I + s * self._onsite(self.R0[i], self.Rx, self.Ry, kx, ky) for i in range(nbl)]
            Su = [s * self.R1[i] for i in range(nbl - 1)]
            Sl = [s * self.R1[i].T for i in range(nbl - 1)]
            S00 = I + s * self._onsite(self.R0c, self.Rx, self.Ry, kx, ky)
            S01r, S01l = s * self.R1c, s * self.R1c.T
            SMCr, SMCl = s * self.R1c, s * self.R1c.T
        c1 = Contact(cblk=nbl, H00=H00, H01=self.A1c, HMC=self.A1c, S00=S00, S01=S01r, SMC=SMCr)
        c2 = Contact(cblk=1, H00=H00, H01=self.A1c.T, HMC=self.A1c.T, S00=S00, S01=S01l,
                     SMC=SMCl)
        return BlockStructure(Hd=Hd, Hu=Hu, Hl=Hl, contacts=[c1, c2], Sd=Sd, Su=Su, Sl=Sl,
                              name=f"kslab n={n} nbl={nbl} k=({kx:.4f},{ky:.4f})")


def config3(n: Optional[int] = None, nbl: Optional[int] = None, with_overlap: bool = True):
    c = C3
    slab = KSlab.make(n or c["n"], nbl or c["nbl"], c["seed"], with_overlap)
    return slab, mp_grid(c["kgrid"]), energy_grid(c["nE"], c["Emin"], c["Emax"]), c["eta"]


def serving_spec(seed: int, n_choices=(32, 64, 128, 256, 512), nbl_range=(10, 200),
                 nE_range=(50, 2000)):
    """C5 request shape from its seed: (n, nbl, nE, kx, ky) -- the same draws as
    ``serving_request``, without building the blocks (the serving loop
    generates them in HBM from the seed)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    n = int(rng.choice(n_choices))
    nbl = int(rng.integers(nbl_range[0], nbl_range[1] + 1))
    nE = int(rng.integers(nE_range[0], nE_range[1] + 1))
    kx, ky = rng.uniform(-np.pi, np.pi, 2)
    return n, nbl, nE, float(kx), float(ky)


def serving_request(seed: int, n_choices=(32, 64, 128, 256, 512), nbl_range=(10, 200),
                    nE_range=(50, 2000)):
    """C5: one request = (KSlab at a single k, energy grid), from a per-request seed."""
    rng = np.random.Generator(np.random.PCG64(seed))
    n = int(rng.choice(n_choices))
    nbl = int(rng.integers(nbl_range[0], nbl_range[1] + 1))
    nE = int(rng.integers(nE_range[0], nE_range[1] + 1))
    slab = KSlab.make(n, nbl, seed=10_000 + seed)
    kx, ky = rng.uniform(-np.pi, np.pi, 2)
    return slab.at_k(kx, ky), energy_grid(nE, -2.0, 2.0), 1e-5
