"""Host-side description of a block-tridiagonal device with semi-infinite contacts.

This is the data model of the reference after ``extract_cont`` and
``zcsr2blk_sod`` (``ln_extract.F90:41-76``, ``sparsekit_drv.F90:3409-3436``):
dense column-major complex128 blocks ``H(i,i)``, ``H(i,i+1)``, ``H(i+1,i)``
(and the same for ``S``) for the ``nbl`` principal layers (PLs), plus for each
contact its first two PLs (``H00``, ``H01``: PL1 and the PL1->PL2 coupling,
pointing away from the device) and the device-contact coupling
``HMC`` = H(device block cblk, contact PL1).  ``cblk`` is 1-based as in the
reference (``ln_structure.F90:80-150``).

Contacts without a surface block only (``surfstart = surfend + 1``), which is
what every configuration here uses.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np


def _f(a) -> np.ndarray:
    return np.asfortranarray(np.asarray(a, dtype=np.complex128))


@dataclass
class Contact:
    cblk: int                       # 1-based device block the contact attaches to
    H00: np.ndarray                 # nc x nc
    H01: np.ndarray                 # nc x nc, PL1 -> PL2 (away from the device)
    HMC: np.ndarray                 # nd x (ns + nc), device block cblk -> surface + PL1
    S00: Optional[np.ndarray] = None
    S01: Optional[np.ndarray] = None
    SMC: Optional[np.ndarray] = None
    # optional surface block of ns orbitals between the device and PL1
    # (surface_green, contselfenergy.F90:186-239)
    Hss: Optional[np.ndarray] = None    # ns x ns
    Hs1: Optional[np.ndarray] = None    # ns x nc, surface -> PL1
    H1s: Optional[np.ndarray] = None    # nc x ns, PL1 -> surface
    Sss: Optional[np.ndarray] = None
    Ss1: Optional[np.ndarray] = None
    S1s: Optional[np.ndarray] = None

    def __post_init__(self):
        self.H00, self.H01, self.HMC = _f(self.H00), _f(self.H01), _f(self.HMC)
        for name in ("S00", "S01", "SMC", "Hss", "Hs1", "H1s", "Sss", "Ss1", "S1s"):
            if getattr(self, name) is not None:
                setattr(self, name, _f(getattr(self, name)))
        if self.Hss is not None and (self.Hs1 is None or self.H1s is None):
            raise ValueError("a surface block needs Hss, Hs1 and H1s")

    @property
    def nc(self) -> int:
        return self.H00.shape[0]

    @property
    def ns(self) -> int:
        return 0 if self.Hss is None else self.Hss.shape[0]


@dataclass
class BlockStructure:
    Hd: List[np.ndarray]
    Hu: List[np.ndarray]
    Hl: List[np.ndarray]
    contacts: List[Contact]
    Sd: Optional[List[np.ndarray]] = None   # None -> identity
    Su: Optional[List[np.ndarray]] = None   # None -> zero
    Sl: Optional[List[np.ndarray]] = None
    name: str = ""

    def __post_init__(self):
        self.Hd = [_f(a) for a in self.Hd]
        self.Hu = [_f(a) for a in self.Hu]
        self.Hl = [_f(a) for a in self.Hl]
        if self.Sd is not None:
            self.Sd = [_f(a) for a in self.Sd]
        if self.Su is not None:
            self.Su = [_f(a) for a in self.Su]
        if self.Sl is not None:
            self.Sl = [_f(a) for a in self.Sl]

    @property
    def nbl(self) -> int:
        return len(self.Hd)

    @property
    def bsz(self) -> List[int]:
        return [a.shape[0] for a in self.Hd]

    @property
    def ndev(self) -> int:
        return int(sum(self.bsz))

    @property
    def ncont(self) -> int:
        return len(self.contacts)

    @property
    def s_identity(self) -> bool:
        return self.Sd is None and self.Su is None and self.Sl is None

    # ---- dense helpers (tests / brute force only) ---------------------------
    def dense_device(self, E: complex) -> np.ndarray:
        """E*S - H of the device region as one dense matrix (no self-energies)."""
        b = self.bsz
        off = np.concatenate([[0], np.cumsum(b)])
        n = off[-1]
        A = np.zeros((n, n), dtype=np.complex128)
        for i in range(self.nbl):
            s = self.Sd[i] if self.Sd is not None else np.eye(b[i])
            A[off[i]:off[i + 1], off[i]:off[i + 1]] = E * s - self.Hd[i]
        for i in range(self.nbl - 1):
            su = self.Su[i] if self.Su is not None else 0.0
            sl = self.Sl[i] if self.Sl is not None else 0.0
            A[off[i]:off[i + 1], off[i + 1]:off[i + 2]] = E * su - self.Hu[i]
            A[off[i + 1]:off[i + 2], off[i]:off[i + 1]] = E * sl - self.Hl[i]
        return A

    def block_offsets(self) -> np.ndarray:
        return np.concatenate([[0], np.cumsum(self.bsz)]).astype(np.int64)
