"""Test-only helpers for the inelastic electron-phonon SCBA: a small k-resolved
chain system and a dense brute-force restatement of the outer SCBA loop
(whole-device inverses instead of the block recursion)."""
import numpy as np

from tests.chains import chain_dense, from_dense


def fermi(E, mu, kT):
    return 1.0 / (1.0 + np.exp((E - mu) / kT))


def bose(w, kT):
    """bose_r, distributions.F90:114-129."""
    if kT == 0.0:
        return 0.0
    return np.exp(-w / kT) if abs(w / kT) > 30.0 else 1.0 / (np.exp(w / kT) - 1.0)


def k_system(nbl=3, n=2, nk=2, seed=3, eps_k=0.3):
    """Chain device of nbl PLs of n sites with chain leads (PL size n); at
    k-point j every site (device and leads) is shifted by eps_k cos(2 pi kx_j)
    on top of a fixed device disorder.  Returns (structs[nk], kfrac[nk,3])."""
    rng = np.random.default_rng(seed)
    nd = nbl * n
    ons = 0.2 * rng.standard_normal(nd)
    kfrac = np.array([[0.1 + 0.25 * j, 0.15 * j, 0.0] for j in range(nk)])
    out = []
    for j in range(nk):
        H = chain_dense(n_dev=nd, n_cont=2 * n, onsite=ons)
        H = H + eps_k * np.cos(2 * np.pi * kfrac[j, 0]) * np.eye(H.shape[0])
        plend = [n * (i + 1) for i in range(nbl)]
        out.append(from_dense(H, None, plend, [nd + 1, nd + 2 * n + 1], [nd + 2 * n, nd + 4 * n], [nbl, 1]))
    return out, kfrac


LATT = np.diag([1.0, 1.0, 10.0])  # lattice vectors as columns


def band_mask(bsz, tridiag):
    """Boolean mask of the blocks an inelastic self-energy covers."""
    off = np.concatenate([[0], np.cumsum(bsz)])
    N = off[-1]
    M = np.zeros((N, N), dtype=bool)
    for i in range(len(bsz)):
        M[off[i]:off[i + 1], off[i]:off[i + 1]] = True
        if tridiag and i + 1 < len(bsz):
            M[off[i]:off[i + 1], off[i + 1]:off[i + 2]] = True
            M[off[i + 1]:off[i + 2], off[i]:off[i + 1]] = True
    return M


def dense_sigmas(Gr, Gn, K, iz, shift, Nq, tridiag, bsz):
    """Sigma_r and Sigma_n of every (k, E) from dense G^r / G^n [nk][ne] on the
    covered blocks: Sigma^<-type sums over q of K(|z_r - z_c|, k, q) times the
    G of the shifted energies (0 outside the grid)."""
    nk, ne = len(Gr), len(Gr[0])
    M = band_mask(bsz, tridiag)
    off = np.concatenate([[0], np.cumsum(bsz)])
    D = np.abs(iz[:, None] - iz[None, :])
    Sr = [[None] * ne for _ in range(nk)]
    Sn = [[None] * ne for _ in range(nk)]
    zero = np.zeros_like(Gr[0][0])
    for k in range(nk):
        for e in range(ne):
            sr = np.zeros_like(zero)
            sn = np.zeros_like(zero)
            for q in range(nk):
                Kq = K[D, k, q]
                grm = Gr[q][e - shift] if e - shift >= 0 else zero
                grp = Gr[q][e + shift] if e + shift < ne else zero
                gnm = Gn[q][e - shift] if e - shift >= 0 else zero
                gnp = Gn[q][e + shift] if e + shift < ne else zero
                sn += Kq * (Nq * gnm + (Nq + 1) * gnp)
                sr += Kq * ((Nq + 1) * grm + Nq * grp + 0.5j * gnm - 0.5j * gnp)
            sr[~M] = 0.0
            sn[~M] = 0.0
            # Sigma_n(i+1,i) is taken as Sigma_n(i,i+1)^H
            for i in range(len(bsz) - 1):
                if tridiag:
                    sn[off[i + 1]:off[i + 2], off[i]:off[i + 1]] = \
                        sn[off[i]:off[i + 1], off[i + 1]:off[i + 2]].conj().T
            Sr[k][e] = sr
            Sn[k][e] = sn
    return Sr, Sn


def dense_inelastic(structs, kw, E, delta, frm, ref, K, iz, shift, Nq, tridiag, niter, tol,
                    sigma_of, g_spin=2.0):
    """The outer inelastic SCBA with dense whole-device inverses (no elastic
    model).  sigma_of(s, c, z) -> contact self-energy on its device block.
    Trapezoidal current integration.  Returns (curr_mat[ne, ncont], currents, passes)."""
    nk, ne = len(structs), len(E)
    s0 = structs[0]
    bsz = list(s0.bsz)
    off = np.concatenate([[0], np.cumsum(bsz)])
    ncont = s0.ncont
    sig = [[[sigma_of(structs[k], c, E[e] + 1j * delta) for c in range(ncont)] for e in range(ne)]
           for k in range(nk)]
    Sr = Sn = None
    Jold = None
    dE = E[1] - E[0]
    w = np.full(ne, dE)
    w[0] = w[-1] = 0.5 * dE
    p = 0
    conv = False
    while not conv and p <= niter:
        Gr = [[None] * ne for _ in range(nk)]
        Gn = [[None] * ne for _ in range(nk)]
        cm = np.zeros((ne, ncont))
        for k in range(nk):
            s = structs[k]
            for e in range(ne):
                A = s.dense_device(E[e])
                Sin = np.zeros_like(A)
                for c, ct in enumerate(s.contacts):
                    b = ct.cblk - 1
                    blk = slice(off[b], off[b + 1])
                    A[blk, blk] -= sig[k][e][c]
                    fd = frm[e][c] - frm[e][ref - 1]
                    if c != ref - 1 and abs(fd) > 0.0:
                        Sg = sig[k][e][c]
                        Sin[blk, blk] += fd * 1j * (Sg - Sg.conj().T)
                if p > 0:
                    A -= Sr[k][e]
                    Sin += Sn[k][e]
                G = np.linalg.inv(A)
                Gnn = G @ Sin @ G.conj().T
                Gr[k][e] = G
                Gn[k][e] = Gnn
                for c, ct in enumerate(s.contacts):
                    b = ct.cblk - 1
                    blk = slice(off[b], off[b + 1])
                    Sg = sig[k][e][c]
                    Gam = 1j * (Sg - Sg.conj().T)
                    Ab = 1j * (G[blk, blk] - G[blk, blk].conj().T)
                    cm[e, c] += -np.real(np.trace(Gam @ Gnn[blk, blk] - frm[e][c] * Gam @ Ab)) * kw[k]
        Sr, Sn = dense_sigmas(Gr, Gn, K, iz, shift, Nq, tridiag, bsz)
        J = np.array([np.sum(cm[:, c] * w) for c in range(ncont)]) * g_spin
        if Jold is not None:
            conv = np.max(np.abs(Jold - J)) < tol
        Jold = J
        p += 1
    return cm, Jold, p
