// hn_fused.hip -- the whole two-contact transmission of one (E,k) item in one
// workgroup, for small blocks.
//
// For principal layers of up to kFusedMaxN orbitals every block fits in LDS
// many times over, and the batched engine's cost is launches and host polls
// (one launch per block operation per step, one convergence poll per
// decimation iteration).  Here one wave per item runs, without leaving the
// kernel: each contact's Sancho-Rubio decimation with its own convergence
// test (contselfenergy.F90:271-340; both stopping rules), the self-energy
// Sigma = Tlc GS Tcl (:445-491, 552-566), the backward sweep at real E
// (build_ESH, iterative.F90:1956-1982; calculate_gsmr_blocks,
// iterative_cpu.F90:51-124), G11 (:172-198) and T = |Re Tr[Gam A - Gam G Gam
// G^dag]| (:399-470).  Block inverses are ZGETRF (row interchanges, pivot =
// first largest |Re|+|Im| as IZAMAX) + ZGETRS against the identity
// (inversions.F90:807-839).  One launch covers every item of a call.
#include "hn_engine.h"
#include "hn_prof.h"

namespace hn {

namespace {

typedef double2 C;

__device__ inline C cmul_(C a, C b) { return C{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
__device__ inline double cabs1_(C a) { return fabs(a.x) + fabs(a.y); }
// a / b, Smith's algorithm; 0 for b == 0
__device__ inline C cdiv_(C a, C b) {
  if (b.x == 0.0 && b.y == 0.0) return C{0.0, 0.0};
  if (fabs(b.x) >= fabs(b.y)) {
    const double r = b.y / b.x, d = b.x + r * b.y;
    return C{(a.x + a.y * r) / d, (a.y - a.x * r) / d};
  }
  const double r = b.x / b.y, d = b.y + r * b.x;
  return C{(a.x * r + a.y) / d, (a.y * r - a.x) / d};
}

// Offsets of an item's blocks in its structure slot, in the table built by
// the host: per PL i: [HD, SD], per interface i: [HU, SU, HL, SL], per
// contact c: [H00, S00, H01, S01, HMC, SMC]; -1: absent.
struct FusedLayout {
  int n, nbl;
  int cblk[2];
  int ct1;            // contact whose Gamma enters the trace (the one on block 1)
  const long long *offs;
  long long slot_stride;
  const C *pool;
  __device__ long long pl(int i, int k) const { return offs[2 * i + k]; }
  __device__ long long itf(int i, int k) const { return offs[2 * nbl + 4 * i + k]; }
  __device__ long long cnt(int c, int k) const { return offs[2 * nbl + 4 * (nbl - 1) + 6 * c + k]; }
};

// One wave (64 lanes) per item.  Every helper is called by all 64 lanes and
// ends with a barrier, so LDS results are visible to the next call.
struct Wave {
  int n, lane;
  __device__ void sync() const { __syncthreads(); }

  // out = z S - H (S absent: identity if sid, else zero); dag: its conjugate
  // transpose
  __device__ void form(C *out, const C *H, const C *S, int sid, double zr, double zi,
                       bool dag) const {
    for (int e = lane; e < n * n; e += 64) {
      const int r = e % n, c = e / n;
      C s = S ? S[e] : C{(sid && r == c) ? 1.0 : 0.0, 0.0};
      C h = H[e];
      C v = C{zr * s.x - zi * s.y - h.x, zr * s.y + zi * s.x - h.y};
      if (dag) out[c + r * n] = C{v.x, -v.y};
      else out[e] = v;
    }
    sync();
  }
  __device__ void copy(C *dst, const C *src) const {
    for (int e = lane; e < n * n; e += 64) dst[e] = src[e];
    sync();
  }
  // dst += s * src (s = +-1)
  __device__ void axpy(C *dst, const C *src, double s) const {
    for (int e = lane; e < n * n; e += 64) dst[e] = C{dst[e].x + s * src[e].x, dst[e].y + s * src[e].y};
    sync();
  }
  // Cm = A B + beta Cm (beta 0 or 1), Cm distinct from A and B
  __device__ void mm(C *Cm, const C *A, const C *B, int beta) const {
    for (int e = lane; e < n * n; e += 64) {
      const int r = e % n, c = e / n;
      double xr = 0.0, xi = 0.0;
      for (int k = 0; k < n; ++k) {
        const C a = A[r + k * n], b = B[k + c * n];
        xr = fma(a.x, b.x, xr);
        xr = fma(-a.y, b.y, xr);
        xi = fma(a.x, b.y, xi);
        xi = fma(a.y, b.x, xi);
      }
      if (beta) { xr += Cm[e].x; xi += Cm[e].y; }
      Cm[e] = C{xr, xi};
    }
    sync();
  }
  // maximum that propagates NaN (a non-finite iterate must stop the loop)
  __device__ double wave_max(double v) const {
    for (int off = 32; off >= 1; off >>= 1) {
      const double o = __shfl_xor(v, off);
      v = (o > v || o != o) ? o : v;
    }
    return v;
  }
  __device__ double wave_sum(double v) const {
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off);
    return v;
  }
  // X = A^-1 by ZGETRF + ZGETRS(I); LU: n x n scratch, piv: n ints of LDS.
  // Returns nonzero for an exactly zero pivot.
  __device__ int inv(C *X, const C *A, C *LU, int *piv) const {
    copy(LU, A);
    int info = 0;
    for (int k = 0; k < n; ++k) {
      // IZAMAX over rows k..n-1 of column k: first largest |Re| + |Im|
      double v = (lane >= k && lane < n) ? cabs1_(LU[lane + k * n]) : -1.0;
      int idx = lane;
      for (int off = 32; off >= 1; off >>= 1) {
        const double ov = __shfl_xor(v, off);
        const int oi = __shfl_xor(idx, off);
        if (ov > v || (ov == v && oi < idx)) { v = ov; idx = oi; }
      }
      const int p = idx;
      if (lane == 0) piv[k] = p;
      if (!(v > 0.0)) {  // ZGETF2: record the first zero pivot, no elimination
        info = info ? info : k + 1;
        sync();
        continue;
      }
      if (p != k && lane < n) {
        const C t = LU[k + lane * n];
        LU[k + lane * n] = LU[p + lane * n];
        LU[p + lane * n] = t;
      }
      sync();
      const C r = cdiv_(C{1.0, 0.0}, LU[k + k * n]);
      if (lane > k && lane < n) LU[lane + k * n] = cmul_(LU[lane + k * n], r);
      sync();
      const int m = n - k - 1;
      for (int e = lane; e < m * m; e += 64) {
        const int i = k + 1 + e % m, j = k + 1 + e / m;
        const C l = LU[i + k * n], u = LU[k + j * n];
        C x = LU[i + j * n];
        x.x = fma(-l.x, u.x, x.x);
        x.x = fma(l.y, u.y, x.x);
        x.y = fma(-l.x, u.y, x.y);
        x.y = fma(-l.y, u.x, x.y);
        LU[i + j * n] = x;
      }
      sync();
    }
    if (info) return info;
    // ZGETRS against I, column `lane` of X: row interchanges, L y = P e, U x = y
    if (lane < n) {
      C *x = X + lane * n;
      for (int i = 0; i < n; ++i) x[i] = C{i == lane ? 1.0 : 0.0, 0.0};
      for (int k = 0; k < n; ++k) {
        const int p = piv[k];
        if (p != k) { const C t = x[k]; x[k] = x[p]; x[p] = t; }
      }
      for (int j = 0; j < n; ++j) {
        const C xj = x[j];
        for (int i = j + 1; i < n; ++i) {
          const C l = LU[i + j * n];
          x[i] = C{x[i].x - (l.x * xj.x - l.y * xj.y), x[i].y - (l.x * xj.y + l.y * xj.x)};
        }
      }
      for (int j = n - 1; j >= 0; --j) {
        x[j] = cdiv_(x[j], LU[j + j * n]);
        const C xj = x[j];
        for (int i = 0; i < j; ++i) {
          const C u = LU[i + j * n];
          x[i] = C{x[i].x - (u.x * xj.x - u.y * xj.y), x[i].y - (u.x * xj.y + u.y * xj.x)};
        }
      }
    }
    sync();
    return 0;
  }
};

constexpr int kFusedMats = 12;

__global__ __launch_bounds__(64) void k_t_fused(FusedLayout L, int nitems, const double *E,
                                                const int *slot, double delta, double sgfacc,
                                                int crit, int maxit, double *T, int *ncyc,
                                                int *status) {
  const int b = blockIdx.x;
  if (b >= nitems) return;
  extern __shared__ __align__(16) unsigned char fsm[];
  const int n = L.n, nn = n * n;
  C *M[kFusedMats];
  for (int i = 0; i < kFusedMats; ++i) M[i] = (C *)fsm + (size_t)i * nn;
  int *piv = (int *)((C *)fsm + (size_t)kFusedMats * nn);
  Wave w{n, (int)threadIdx.x};
  const C *base = L.pool + (long long)slot[b] * L.slot_stride;
  auto blk = [&](long long off) -> const C * { return off < 0 ? nullptr : base + off; };
  const double er = E[b];
  int st = 0;
  C *Sig[2] = {M[10], M[11]};

  for (int c = 0; c < 2; ++c) {
    C *Ao = M[0], *Aos = M[1], *Bo = M[2], *Co = M[3], *Go = M[4];
    C *GC = M[5], *C1 = M[6], *GB = M[7], *B1 = M[8], *LU = M[9];
    const C *H00 = blk(L.cnt(c, 0)), *S00 = blk(L.cnt(c, 1));
    const C *H01 = blk(L.cnt(c, 2)), *S01 = blk(L.cnt(c, 3));
    w.form(Ao, H00, S00, 1, er, delta, false);
    w.copy(Aos, Ao);
    w.form(Bo, H01, S01, 0, er, delta, false);
    w.form(Co, H01, S01, 0, er, -delta, true);
    int okco = 0, it = 1, nc = 0;
    for (; it <= maxit; ++it) {
      if (w.inv(Go, Ao, LU, piv)) st |= ST_SINGULAR;
      w.mm(GC, Go, Co, 0);
      w.mm(C1, Co, GC, 0);
      double m = 0.0;
      if (crit == 0) {
        for (int e = w.lane; e < nn; e += 64) {
          const double a = hypot(C1[e].x, C1[e].y);
          m = (a > m || a != a) ? a : m;
        }
        m = w.wave_max(m);
      } else {
        for (int e = w.lane; e < nn; e += 64) m += fabs(C1[e].x) + fabs(C1[e].y);
        m = w.wave_sum(m);
      }
      if (m != m) { st |= ST_NONFINITE; nc = it; break; }
      if (m <= sgfacc) {
        if (okco) { nc = it; break; }
        okco = 1;
      } else {
        okco = 0;
      }
      // Self = Bo GoXCo: Ao_s -= Self, Ao -= Self
      w.mm(LU, Bo, GC, 0);
      w.axpy(Aos, LU, -1.0);
      w.axpy(Ao, LU, -1.0);
      w.mm(GB, Go, Bo, 0);
      w.mm(B1, Bo, GB, 0);
      w.mm(LU, Co, GB, 0);
      w.axpy(Ao, LU, -1.0);
      C *t = Bo; Bo = B1; B1 = t;  // Bo = B1
      t = Co; Co = C1; C1 = t;     // Co = C1
      // keep the named buffers consistent for the next iteration
      M[2] = Bo; M[8] = B1; M[3] = Co; M[6] = C1;
    }
    if (it > maxit) { nc = maxit + 1; st |= ST_NOT_CONVERGED; }
    if (w.inv(Go, Aos, LU, piv)) st |= ST_SINGULAR;
    // Sigma = Tlc GS Tcl, Tlc = z SMC - HMC, Tcl = (z* SMC - HMC)^dag
    const C *HMC = blk(L.cnt(c, 4)), *SMC = blk(L.cnt(c, 5));
    w.form(GC, HMC, SMC, 0, er, delta, false);
    w.form(C1, HMC, SMC, 0, er, -delta, true);
    w.mm(GB, GC, Go, 0);
    w.mm(Sig[c], GB, C1, 0);
    if (w.lane == 0 && ncyc) ncyc[(long long)b * 2 + c] = nc;
    // restore the buffer table for the next contact
    for (int i = 0; i < 10; ++i) M[i] = (C *)fsm + (size_t)i * nn;
  }

  // backward sweep at real E: g_nbl = ESH_nbl^-1, g_i = (ESH_ii - ESH_i,i+1 g ESH_i+1,i)^-1
  C *g = M[0], *D = M[1], *U = M[2], *Lw = M[3], *W = M[4], *X = M[5], *LU = M[6];
  auto esh_diag = [&](int i) {
    w.form(D, blk(L.pl(i, 0)), blk(L.pl(i, 1)), 1, er, 0.0, false);
    for (int c = 0; c < 2; ++c)
      if (L.cblk[c] == i + 1) w.axpy(D, Sig[c], -1.0);
  };
  esh_diag(L.nbl - 1);
  if (w.inv(g, D, LU, piv)) st |= ST_SINGULAR;
  for (int i = L.nbl - 2; i >= 0; --i) {
    esh_diag(i);
    w.form(U, blk(L.itf(i, 0)), blk(L.itf(i, 1)), 0, er, 0.0, false);
    w.form(Lw, blk(L.itf(i, 2)), blk(L.itf(i, 3)), 0, er, 0.0, false);
    w.mm(W, U, g, 0);
    w.mm(X, W, Lw, 0);
    w.axpy(D, X, -1.0);
    if (w.inv(g, D, LU, piv)) st |= ST_SINGULAR;
  }
  // g = G11.  T = |Re Tr[Gam A - Gam G Gam G^dag]|, A = i (G - G^dag)
  C *Gam = M[1], *GA = M[2], *W1 = M[3], *W2 = M[4], *AA = M[5];
  const C *Sg = Sig[L.ct1];
  for (int e = w.lane; e < nn; e += 64) {
    const int r = e % n, c = e / n;
    const C s = Sg[e], st2 = Sg[c + r * n];
    Gam[e] = C{-(s.y + st2.y), s.x - st2.x};  // i (S - S^dag)
    const C gt = g[c + r * n];
    GA[e] = C{gt.x, -gt.y};
  }
  w.sync();
  for (int e = w.lane; e < nn; e += 64) {
    const C d = C{g[e].x - GA[e].x, g[e].y - GA[e].y};
    AA[e] = C{-d.y, d.x};
  }
  w.sync();
  w.mm(W1, Gam, g, 0);
  w.mm(W2, W1, Gam, 0);
  w.mm(W1, W2, GA, 0);
  w.mm(W2, Gam, AA, 0);
  double tr = 0.0;
  for (int i = w.lane; i < n; i += 64) tr += W2[i + i * n].x - W1[i + i * n].x;
  tr = w.wave_sum(tr);
  if (w.lane == 0) {
    T[b] = fabs(tr);
    status[b] = st;
  }
}

}  // namespace

bool Engine::t_fused(DevStruct &st, int nitems, const double *dE, const int *dslot,
                     const SolveOpts &o, int ni, int nf, double *dT, int *dncyc, int *dstatus) {
  if (st.ncont != 2 || !st.uniform() || st.bsz[0] > kFusedMaxN || st.tun_mask) return false;
  if (st.ns[0] || st.ns[1] || st.fict[0] || st.fict[1]) return false;
  if (st.nc[0] != st.bsz[0] || st.nc[1] != st.bsz[0]) return false;
  if (o.precision != PREC_FP64 || o.sgf_mode != 0 || o.dEi) return false;
  if (std::min(st.cblk[ni - 1], st.cblk[nf - 1]) != 1) return false;
  const int n = st.bsz[0], nbl = st.nbl;
  std::vector<long long> offs;
  auto off = [&](int kind, int i) -> long long {
    return st.off[kind].empty() ? -1 : st.off[kind][i];
  };
  for (int i = 0; i < nbl; ++i) { offs.push_back(off(BK_HD, i)); offs.push_back(off(BK_SD, i)); }
  for (int i = 0; i + 1 < nbl; ++i) {
    offs.push_back(off(BK_HU, i)); offs.push_back(off(BK_SU, i));
    offs.push_back(off(BK_HL, i)); offs.push_back(off(BK_SL, i));
  }
  for (int c = 0; c < 2; ++c)
    for (int k : {BK_H00, BK_S00, BK_H01, BK_S01, BK_HMC, BK_SMC}) offs.push_back(off(k, c));
  arena_.reserve(offs.size() * sizeof(long long) + 4096);
  arena_.reset();
  long long *doffs = arena_.take<long long>(offs.size());
  HN_CHECK(hipMemcpyAsync(doffs, offs.data(), offs.size() * sizeof(long long),
                          hipMemcpyHostToDevice, stream_));
  FusedLayout L;
  L.n = n;
  L.nbl = nbl;
  L.cblk[0] = st.cblk[0];
  L.cblk[1] = st.cblk[1];
  const int a = ni - 1, bb = nf - 1;
  L.ct1 = st.cblk[a] < st.cblk[bb] ? a : bb;
  L.offs = doffs;
  L.slot_stride = st.slot_elems;
  L.pool = st.pool;
  const size_t lds = (size_t)kFusedMats * n * n * sizeof(double2) + 64 * sizeof(int);
  {
    ProfScope ps(PT_FUSED, -1.0, 0.0, stream_);
    hipLaunchKernelGGL(k_t_fused, dim3(nitems), dim3(64), lds, stream_, L, nitems, dE, dslot,
                       o.delta, o.sgfacc, o.dec_crit, o.dec_maxit, dT, dncyc, dstatus);
    HN_CHECK(hipGetLastError());
  }
  HN_CHECK(hipStreamSynchronize(stream_));  // the offset table's host copy stays valid until here
  return true;
}

}  // namespace hn
