"""Small test structures (test-only helpers)."""
import numpy as np

from libnegf_amd.structure import BlockStructure, Contact


def chain_structure(npl=6, plsz=10, t=1.0, onsite=None):
    """Reconstruction of the reference tests' 1-D chain: 60-site device in 6
    PLs of 10 sites, contact 1 (sites 61..80) on block 6, contact 2 (sites
    81..100) on block 1, hopping t, S = I (tests/f90elph_deph/hello.F90:42-48,
    tests/c_int_elph_deph/c_int_elph_deph.c:15-21).  The reference's H data
    are git-LFS objects; this rebuilds an equivalent uniform chain."""
    H00 = np.zeros((plsz, plsz))
    for i in range(plsz - 1):
        H00[i, i + 1] = H00[i + 1, i] = t
    Hd = []
    for p in range(npl):
        h = H00.copy()
        if onsite is not None:
            h += np.diag(onsite[p * plsz:(p + 1) * plsz])
        Hd.append(h)
    U = np.zeros((plsz, plsz))
    U[plsz - 1, 0] = t
    Hu = [U.copy() for _ in range(npl - 1)]
    Hl = [U.T.copy() for _ in range(npl - 1)]
    H01 = U.copy()               # PL1 -> PL2 within a contact (last site -> first site)
    HMC_r = np.zeros((plsz, plsz)); HMC_r[plsz - 1, 0] = t   # site 60 <-> 61
    HMC_l = np.zeros((plsz, plsz)); HMC_l[0, 0] = t          # site 1  <-> 81
    c1 = Contact(cblk=npl, H00=H00, H01=H01, HMC=HMC_r)
    c2 = Contact(cblk=1, H00=H00, H01=H01, HMC=HMC_l)
    return BlockStructure(Hd=Hd, Hu=Hu, Hl=Hl, contacts=[c1, c2], name="chain")


def dense_G(s, E, sigmas):
    """Brute force: G = (E S - H - sum Sigma)^-1 of the whole device (numpy/LAPACK)."""
    A = s.dense_device(E)
    off = s.block_offsets()
    for c, Sg in zip(s.contacts, sigmas):
        b = c.cblk - 1
        A[off[b]:off[b + 1], off[b]:off[b + 1]] -= Sg
    return np.linalg.inv(A), off


def gamma(S):
    return 1j * (S - S.conj().T)


def brute_T(s, E, sigmas, ci=0, cf=1):
    G, off = dense_G(s, E, sigmas)
    bi, bf = s.contacts[ci].cblk - 1, s.contacts[cf].cblk - 1
    Gfi = G[off[bf]:off[bf + 1], off[bi]:off[bi + 1]]
    return float(np.real(np.trace(gamma(sigmas[cf]) @ Gfi @ gamma(sigmas[ci]) @ Gfi.conj().T)))


def from_dense(H, S, plend, surfstart, contend, cblk, surfend=None):
    """BlockStructure of a dense PL-ordered H (and S, or None for identity) with
    the reference's structure description (1-based): device PLs end at plend,
    contact c occupies surfstart[c]..contend[c] -- a surface block
    surfstart..surfend (empty by default, surfend = surfstart - 1) and two
    principal layers -- and couples to PL cblk[c] through the surface and the
    first PL (ln_extract.F90:41-76)."""
    pls = [0] + [int(p) for p in plend[:-1]]
    ple = [int(p) for p in plend]
    blk = lambda M, a, b, c, d: np.array(M[a:b, c:d], dtype=np.complex128)
    nbl = len(ple)
    Hd = [blk(H, pls[i], ple[i], pls[i], ple[i]) for i in range(nbl)]
    Hu = [blk(H, pls[i], ple[i], pls[i + 1], ple[i + 1]) for i in range(nbl - 1)]
    Hl = [blk(H, pls[i + 1], ple[i + 1], pls[i], ple[i]) for i in range(nbl - 1)]
    Sd = Su = Sl = None
    if S is not None:
        Sd = [blk(S, pls[i], ple[i], pls[i], ple[i]) for i in range(nbl)]
        Su = [blk(S, pls[i], ple[i], pls[i + 1], ple[i + 1]) for i in range(nbl - 1)]
        Sl = [blk(S, pls[i + 1], ple[i + 1], pls[i], ple[i]) for i in range(nbl - 1)]
    cs = []
    for c in range(len(cblk)):
        s0 = int(surfstart[c]) - 1                                   # surface start (0-based)
        a = s0 if surfend is None else int(surfend[c])               # first PL start
        n = (int(contend[c]) - a) // 2
        b = int(cblk[c]) - 1
        kw = dict(cblk=int(cblk[c]), H00=blk(H, a, a + n, a, a + n), H01=blk(H, a, a + n, a + n, a + 2 * n),
                  HMC=blk(H, pls[b], ple[b], s0, a + n))
        if a > s0:
            kw.update(Hss=blk(H, s0, a, s0, a), Hs1=blk(H, s0, a, a, a + n), H1s=blk(H, a, a + n, s0, a))
        if S is not None:
            kw.update(S00=blk(S, a, a + n, a, a + n), S01=blk(S, a, a + n, a + n, a + 2 * n),
                      SMC=blk(S, pls[b], ple[b], s0, a + n))
            if a > s0:
                kw.update(Sss=blk(S, s0, a, s0, a), Ss1=blk(S, s0, a, a, a + n), S1s=blk(S, a, a + n, s0, a))
        cs.append(Contact(**kw))
    return BlockStructure(Hd=Hd, Hu=Hu, Hl=Hl, contacts=cs, Sd=Sd, Su=Su, Sl=Sl, name="dense")


def chain_dense(n_dev=60, n_cont=20, t=1.0, onsite=None):
    """Dense 100-site H of the reference tests' chain with the site order of
    tests/c_int_elph_deph/c_int_elph_deph.c:15-21 (device 1..60, contact 1
    61..80 attached at site 60, contact 2 81..100 attached at site 1)."""
    N = n_dev + 2 * n_cont
    H = np.zeros((N, N), dtype=np.complex128)
    for i in range(1, n_dev + n_cont):           # 1..80 chain
        H[i, i - 1] = H[i - 1, i] = t
    for i in range(n_dev + n_cont + 1, N):       # 81..100 chain
        H[i, i - 1] = H[i - 1, i] = t
    H[0, n_dev + n_cont] = H[n_dev + n_cont, 0] = t   # site 1 <-> 81
    if onsite is not None:
        H[np.arange(n_dev), np.arange(n_dev)] += onsite
    return H
