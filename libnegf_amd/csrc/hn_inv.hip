// hn_inv.hip -- batched complex block inverse with partial pivoting.
//
// Replaces zinv (ZGETRF + ZGETRS against I, inversions.F90:807-839) and
// cu_Zinverse (gpuroutines.cu:449-523) for whole batches of blocks.
//
// Small blocks (n^2 complex fit in LDS): one workgroup per block runs an
// in-place Gauss-Jordan elimination with partial pivoting entirely in LDS,
// followed by the column interchanges that undo the row pivoting.
//
// Large blocks: two-level blocked in-place Gauss-Jordan with *lazy* pivoting.
// Rows are never moved; the pivot row of each logical column is recorded.
// Per inner panel K of nb columns (one launch, one workgroup per item, the
// panel in registers): partial-pivoting LU of the panel over the
// not-yet-pivoted rows picks nb pivot rows; with Q = B(piv, K) every thread
// forms its row of Z = [-B(r,K) Q^-1 ; Q^-1 - I], writes the new panel
// columns, applies B(:, J) += Z R(:, J) to the other columns J of the current
// super-panel S (kInvSuper columns; R = their pivot rows, staged in LDS), and
// copies the pivot rows of the columns outside S.
// After the inner panels of S, its columns hold Zhat + E where
// I + Zhat P^T is the product of the inner rank-nb transforms (the last panel
// of S writes Zhat), so the rest of the matrix takes ONE deferred
// rank-kInvSuper update
//   B(:, J) += Zhat R(:, J),  J outside S  (R = pivot rows before S started),
// a GEMM on the matrix cores.
// After the last panel, Ainv(i, c) = B(pi(i), pi^-1(c)) with pi(c) the pivot
// row of logical column c.
#include "hn_gemm.h"
#include "hn_inv.h"
#include "hn_prof.h"

namespace hn {

namespace {

constexpr int NT = 256;
constexpr int kLdsBudget = 160 * 1024 - 2048;

template <typename R> struct CT_;
template <> struct CT_<double> { typedef double2 t; };
template <> struct CT_<float> { typedef float2 t; };

template <typename C> __device__ inline C cmul(C a, C b) {
  return C{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename C> __device__ inline C csub(C a, C b) { return C{a.x - b.x, a.y - b.y}; }
template <typename C> __device__ inline C cneg(C a) { return C{-a.x, -a.y}; }
template <typename C> __device__ inline C crecip(C b) {
  // 1/b, Smith's algorithm; 0 for b == 0 (singular: flagged by the caller)
  typedef decltype(b.x) R;
  if (b.x == R(0) && b.y == R(0)) return C{0, 0};
  if (fabs(b.x) >= fabs(b.y)) {
    R r = b.y / b.x, d = b.x + r * b.y;
    return C{R(1) / d, -r / d};
  } else {
    R r = b.x / b.y, d = b.y + r * b.x;
    return C{r / d, R(-1) / d};
  }
}
template <typename C> __device__ inline auto cabs1(C a) -> decltype(a.x) {
  return fabs(a.x) + fabs(a.y);
}

// wave-level argmax of (v, i): larger v wins, ties -> smaller index
template <typename R> __device__ inline void wave_argmax(R &v, int &i) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    R ov = __shfl_xor(v, off);
    int oi = __shfl_xor(i, off);
    if (ov > v || (ov == v && oi >= 0 && (i < 0 || oi < i))) {
      v = ov;
      i = oi;
    }
  }
}

// ---------------------------------------------------------------------------
// single-workgroup LDS Gauss-Jordan
// ---------------------------------------------------------------------------
template <typename R>
__global__ __launch_bounds__(NT) void k_inv_lds(int n, typename CT_<R>::t *A, long long sA,
                                                typename CT_<R>::t *Ainv, long long sAinv,
                                                const int *active, int *status) {
  typedef typename CT_<R>::t C;
  extern __shared__ __align__(16) unsigned char smem[];
  const int b = blockIdx.x;
  if (active && active[b] == 0) return;
  C *M = (C *)smem;
  C *rowbuf = M + (size_t)n * n;
  C *colbuf = rowbuf + n;
  int *perm = (int *)(colbuf + n);
  __shared__ int s_p, s_sing;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const C *Ab = A + (long long)b * sA;
  for (int e = tid; e < n * n; e += NT) M[e] = Ab[e];
  if (tid == 0) s_sing = 0;
  // 2-D thread mapping for the rank-1 update: thread -> (row i, column phase)
  const int cpar = NT / n;  // n <= NT here
  const int ui = tid % n, uj0 = tid / n;
  const bool upd = uj0 < cpar;
  __syncthreads();
  for (int k = 0; k < n; ++k) {
    if (wave == 0) {
      R best = R(-1);
      int bi = -1;
      for (int i = k + lane; i < n; i += 64) {
        R v = cabs1(M[i + (size_t)k * n]);
        if (v > best) { best = v; bi = i; }
      }
      wave_argmax(best, bi);
      if (lane == 0) {
        s_p = bi < 0 ? k : bi;
        if (!(best > R(0))) s_sing = 1;
      }
    }
    __syncthreads();
    const int p = s_p;
    if (p != k)
      for (int j = tid; j < n; j += NT) {
        C t = M[k + (size_t)j * n];
        M[k + (size_t)j * n] = M[p + (size_t)j * n];
        M[p + (size_t)j * n] = t;
      }
    if (tid == 0) perm[k] = p;
    __syncthreads();
    const C pinv = crecip(M[k + (size_t)k * n]);
    for (int j = tid; j < n; j += NT) rowbuf[j] = (j == k) ? pinv : cmul(M[k + (size_t)j * n], pinv);
    for (int i = tid; i < n; i += NT) colbuf[i] = (i == k) ? C{0, 0} : M[i + (size_t)k * n];
    __syncthreads();
    if (upd) {
      const C ci = colbuf[ui];
      for (int j = uj0; j < n; j += cpar) {
        size_t e = ui + (size_t)j * n;
        if (ui == k) M[e] = rowbuf[j];
        else if (j == k) M[e] = cneg(cmul(ci, pinv));
        else M[e] = csub(M[e], cmul(ci, rowbuf[j]));
      }
    }
    __syncthreads();
  }
  for (int k = n - 1; k >= 0; --k) {
    const int p = perm[k];
    if (p != k)
      for (int i = tid; i < n; i += NT) {
        C t = M[i + (size_t)k * n];
        M[i + (size_t)k * n] = M[i + (size_t)p * n];
        M[i + (size_t)p * n] = t;
      }
    __syncthreads();
  }
  C *Ob = Ainv + (long long)b * sAinv;
  for (int e = tid; e < n * n; e += NT) Ob[e] = M[e];
  if (tid == 0 && s_sing && status) atomicOr(&status[b], (int)ST_SINGULAR);
}

// ---------------------------------------------------------------------------
// blocked path
// ---------------------------------------------------------------------------
template <typename R>
__global__ __launch_bounds__(NT) void k_inv_reset(int n, int batch, int *used,
                                                  const int *active) {
  long long e = (long long)blockIdx.x * NT + threadIdx.x;
  if (e >= (long long)n * batch) return;
  int b = (int)(e / n);
  if (active && active[b] == 0) return;
  used[e] = 0;
}

// Register-resident panel kernel: thread t owns rows t + NT*q (q < RPT) of the
// n x nb panel in registers.  Column j: block argmax over unused rows (one
// wave shuffle reduction + one LDS exchange), the owner of the pivot row
// publishes its row, every thread eliminates its own rows (multipliers kept in
// place, LU storage).  Two barriers per column.  Afterwards the pivot block is
// Q = L U (rows in selection order) and each thread solves W Q = B(r, K) for
// its rows by one forward and one backward substitution (pivot rows solve
// with e_j, giving row j of Q^-1), so Q is never inverted explicitly.
//
// The same kernel applies the panel's transform to the other columns of the
// current super-panel [s0, s0+sw): B(:, J) += Z R(:, J) with R = the pivot
// rows of those columns (staged in LDS before any row is updated) and Z held
// in registers, one row per thread -- so no separate rank-nb GEMM pass runs
// over the super-panel.
template <typename R, int NB, int RPT>
__global__ __launch_bounds__(NT) void k_inv_panel_reg(int n, int c0, int nbp, int rb_row, int s0,
                                                      int sw,
                                                      typename CT_<R>::t *__restrict__ A,
                                                      long long sA,
                                                      typename CT_<R>::t *__restrict__ Z,
                                                      typename CT_<R>::t *__restrict__ Zh,
                                                      typename CT_<R>::t *__restrict__ Rb,
                                                      int *__restrict__ used_g,
                                                      int *__restrict__ piv_g,
                                                      const int *__restrict__ active,
                                                      int *__restrict__ status) {
  typedef typename CT_<R>::t C;
  const int b = blockIdx.x;
  if (active && active[b] == 0) return;
  __shared__ C LU[NB][NB];  // LU[j][l]: row j of the pivot block's LU factors
  __shared__ C Rs[NB][kInvSuper];  // pivot rows of the other super-panel columns
  __shared__ int spiv[kInvSuper];  // pivot rows of the super-panel's earlier columns
  __shared__ int piv[NB];
  __shared__ R redv[NT / 64];
  __shared__ int redi[NT / 64];
  __shared__ int s_sing;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  C *Ab = A + (long long)b * sA;
  C *Zb = Z + (long long)b * n * kInvPanel;
  C *Rbb = Rb + (long long)b * kInvSuper * n + rb_row;
  int *ug = used_g + (long long)b * n;

  C P[RPT][NB];
  int used[RPT], pos[RPT];
#pragma unroll
  for (int q = 0; q < RPT; ++q) {
    const int r = tid + NT * q;
    const bool valid = r < n;
    used[q] = valid ? ug[r] : 1;
    pos[q] = -1;
#pragma unroll
    for (int j = 0; j < NB; ++j)
      P[q][j] = (valid && j < nbp) ? Ab[r + (size_t)(c0 + j) * n] : C{0, 0};
  }
  if (tid == 0) s_sing = 0;

#pragma unroll
  for (int j = 0; j < NB; ++j) {
    if (j < nbp) {
    // ---- argmax of |.|_1 over unused rows (NaN ranks below every number) ----
    R best = R(-1);
    int bi = -1;
#pragma unroll
    for (int q = 0; q < RPT; ++q)
      if (!used[q]) {
        R v = cabs1(P[q][j]);
        if (!(v == v)) v = R(-0.5);
        if (v > best) { best = v; bi = tid + NT * q; }
      }
    wave_argmax(best, bi);
    if (lane == 0) { redv[wave] = best; redi[wave] = bi; }
    __syncthreads();
    R bv = redv[0];
    int p = redi[0];
#pragma unroll
    for (int w = 1; w < NT / 64; ++w)
      if (redv[w] > bv || (redv[w] == bv && redi[w] >= 0 && (p < 0 || redi[w] < p))) {
        bv = redv[w];
        p = redi[w];
      }
    if (tid == 0 && !(bv > R(0))) s_sing = 1;
    // ---- the owner publishes its (L|U) row ----
    if (p >= 0 && (p % NT) == tid) {
#pragma unroll
      for (int q = 0; q < RPT; ++q)
        if (q == p / NT) {
          used[q] = 1;
          pos[q] = j;
#pragma unroll
          for (int l = 0; l < NB; ++l) LU[j][l] = P[q][l];
        }
      piv[j] = p;
    }
    __syncthreads();
    // ---- eliminate column j from the remaining rows ----
    const C pinv = crecip(LU[j][j]);
#pragma unroll
    for (int q = 0; q < RPT; ++q)
      if (!used[q]) {
        const C l = cmul(P[q][j], pinv);
        P[q][j] = l;
#pragma unroll
        for (int jj = j + 1; jj < NB; ++jj)
          if (jj < nbp) P[q][jj] = csub(P[q][jj], cmul(l, LU[j][jj]));
      }
    }
  }
  __syncthreads();

  // ---- copy of the pivot rows outside the panel columns.  Element e is
  // (pivot j = e % nbp, column c = e / nbp): a wave reads the pivot rows of a
  // few adjacent columns.  Each thread issues all its loads before storing. ----
  {
    constexpr int kPer = 16;  // elements per thread per round
    const int total = nbp * n;
    for (int base = 0; base < total; base += NT * kPer) {
      C v[kPer];
#pragma unroll
      for (int q = 0; q < kPer; ++q) {
        const int e = base + tid + q * NT;
        const int j = e % nbp, c = e / nbp;
        v[q] = (e < total) ? Ab[piv[j] + (size_t)c * n] : C{0, 0};
      }
#pragma unroll
      for (int q = 0; q < kPer; ++q) {
        const int e = base + tid + q * NT;
        const int j = e % nbp, c = e / nbp;
        if (e < total && (c < c0 || c >= c0 + nbp)) {
          if (c >= s0 && c < s0 + sw) Rs[j][c - s0] = v[q];
          else Rbb[j + (size_t)c * kInvSuper] = v[q];
        }
      }
    }
  }
  // reciprocals of the pivots U(l,l), once
  __shared__ C dinv[NB];
  if (tid < nbp) dinv[tid] = crecip(LU[tid][tid]);
  // The last panel of a super-panel also writes Zhat = B(:, S) - E for the
  // deferred update of the other columns (E: ones at (pivot row, column)).
  const bool last = Zh != nullptr && c0 + nbp == s0 + sw;
  if (last && tid < c0 - s0) spiv[tid] = piv_g[(long long)b * n + s0 + tid];
  __syncthreads();
  C *Zhb = last ? Zh + (long long)b * n * kInvSuper : nullptr;

  // ---- W Q = x  with Q = L U:  y U = x (forward), W L = y (backward) ----
#pragma unroll
  for (int q = 0; q < RPT; ++q) {
    const int r = tid + NT * q;
    if (r < n) {
    const int pj = pos[q];
    // the eliminated row is dead now: the right-hand side reuses its registers
#pragma unroll
    for (int l = 0; l < NB; ++l) {
      if (pj >= 0) P[q][l] = C{(l == pj) ? R(1) : R(0), R(0)};
      else P[q][l] = (l < nbp) ? Ab[r + (size_t)(c0 + l) * n] : C{0, 0};
    }
#pragma unroll
    for (int l = 0; l < NB; ++l) {
      if (l < nbp) {
        C acc = P[q][l];
#pragma unroll
        for (int m = 0; m < l; ++m) acc = csub(acc, cmul(P[q][m], LU[m][l]));
        P[q][l] = cmul(acc, dinv[l]);
      }
    }
#pragma unroll
    for (int l = NB - 1; l >= 0; --l) {
      if (l < nbp) {
        C acc = P[q][l];
#pragma unroll
        for (int m = l + 1; m < NB; ++m)
          if (m < nbp) acc = csub(acc, cmul(P[q][m], LU[m][l]));
        P[q][l] = acc;
      }
    }
#pragma unroll
    for (int l = 0; l < NB; ++l) {
      if (l >= nbp) continue;
      C w = P[q][l];
      if (pj >= 0) {
        Ab[r + (size_t)(c0 + l) * n] = w;
        if (l == pj) w.x -= R(1);
      } else {
        w = cneg(w);
        Ab[r + (size_t)(c0 + l) * n] = w;
      }
      P[q][l] = w;  // row r of Z
      if (sw == 0) Zb[r + (size_t)l * n] = w;
      if (last) Zhb[r + (size_t)(c0 - s0 + l) * n] = w;
    }
    // ---- B(r, J) += Z(r, :) R(:, J) for the other super-panel columns J ----
    for (int cb = s0; cb < s0 + sw; cb += 8) {
      C x[8];
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        const int c = cb + k;
        const bool on = c < s0 + sw && (c < c0 || c >= c0 + nbp);
        x[k] = on ? Ab[r + (size_t)c * n] : C{0, 0};
      }
#pragma unroll
      for (int l = 0; l < NB; ++l) {
        if (l < nbp) {
          const C z = P[q][l];
#pragma unroll
          for (int k = 0; k < 8; ++k) {
            const C rv = Rs[l][(cb + k - s0) & (kInvSuper - 1)];
            x[k].x += z.x * rv.x - z.y * rv.y;
            x[k].y += z.x * rv.y + z.y * rv.x;
          }
        }
      }
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        const int c = cb + k;
        if (c < s0 + sw && (c < c0 || c >= c0 + nbp)) {
          Ab[r + (size_t)c * n] = x[k];
          if (last) {  // only earlier columns of S remain here (the panel is the last one)
            C zh = x[k];
            if (spiv[c - s0] == r) zh.x -= R(1);
            Zhb[r + (size_t)(c - s0) * n] = zh;
          }
        }
      }
    }
    ug[r] = used[q];
    }
  }
  if (tid < nbp) piv_g[(long long)b * n + c0 + tid] = piv[tid];
  if (tid == 0 && s_sing && status) atomicOr(&status[b], (int)ST_SINGULAR);
}

template <typename R>
__global__ __launch_bounds__(NT) void k_inv_unperm(int n, const typename CT_<R>::t *B,
                                                   long long sB, typename CT_<R>::t *Out,
                                                   long long sOut, const int *piv_g,
                                                   const int *active) {
  typedef typename CT_<R>::t C;
  extern __shared__ __align__(16) unsigned char smem[];
  const int b = blockIdx.y;
  if (active && active[b] == 0) return;
  int *pv = (int *)smem;
  int *ip = pv + n;
  const int *pg = piv_g + (long long)b * n;
  for (int c = threadIdx.x; c < n; c += NT) {
    int p = pg[c];
    pv[c] = (p >= 0 && p < n) ? p : 0;  // pivots are a permutation of [0, n)
    ip[c] = 0;
  }
  __syncthreads();
  for (int c = threadIdx.x; c < n; c += NT) ip[pv[c]] = c;
  __syncthreads();
  const C *Bb = B + (long long)b * sB;
  C *Ob = Out + (long long)b * sOut;
  // columns handled by this workgroup: blockIdx.x * 8 .. +8
  for (int cc = 0; cc < 8; ++cc) {
    int c = blockIdx.x * 8 + cc;
    if (c >= n) break;
    const int src = ip[c];
    for (int i = threadIdx.x; i < n; i += NT)
      Ob[i + (size_t)c * n] = Bb[pv[i] + (size_t)src * n];
  }
}

template <typename R>
size_t lds_bytes_small(int n) {
  typedef typename CT_<R>::t C;
  return (size_t)n * n * sizeof(C) + 2 * (size_t)n * sizeof(C) + (size_t)n * sizeof(int);
}

template <typename R>
int lds_max_n() {
  int n = 1;
  while (n + 1 <= NT && lds_bytes_small<R>(n + 1) <= (size_t)kLdsBudget) ++n;
  return n;
}

template <typename Kern>
void allow_lds(Kern k, size_t bytes) {
  if (bytes > 64 * 1024)
    HN_CHECK(hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)bytes));
}

template <typename R>
void inv_impl(int n, typename CT_<R>::t *A, long long sA, typename CT_<R>::t *Ainv,
              long long sAinv, int batch, const int *active, int *status, const InvWork<R> &w,
              hipStream_t s) {
  if (n <= 0 || batch <= 0) return;
  if (n <= lds_max_n<R>()) {
    size_t lds = lds_bytes_small<R>(n);
    allow_lds(k_inv_lds<R>, lds);
    ProfScope ps(PT_INV_SMALL, -1.0, 0.0, s);
    hipLaunchKernelGGL(k_inv_lds<R>, dim3(batch), dim3(NT), lds, s, n, A, sA, Ainv, sAinv,
                       active, status);
    HN_CHECK(hipGetLastError());
    return;
  }
  typedef typename CT_<R>::t C;
  (void)sizeof(C);
  if (n > 512) throw std::runtime_error("inverse: block order > 512 is not supported");
  if (sA < (long long)n * n || sAinv < (long long)n * n)
    throw std::runtime_error("inverse: item stride smaller than n*n");
  if (!w.Z || w.n < n || w.batch < batch) throw std::runtime_error("inverse: workspace too small");
  // panel width: FP32 32 up to n = 256; FP64 16 (a 32-wide FP64 panel row
  // exceeds the register file); rows per thread RPT = 1 up to n = 256, else 2
  const int NB = (sizeof(R) == 4 && n <= NT) ? 32 : 16;
  long long nb_tot = (long long)n * batch;
  hipLaunchKernelGGL(k_inv_reset<R>, dim3((unsigned)((nb_tot + NT - 1) / NT)), dim3(NT), 0, s,
                     n, batch, w.used, active);
  for (int s0 = 0; s0 < n; s0 += kInvSuper) {
    const int sw = std::min(kInvSuper, n - s0);
    for (int c0 = s0; c0 < s0 + sw; c0 += NB) {
      const int nbp = std::min(NB, s0 + sw - c0);
      {
        ProfScope ps(PT_INV_PANEL, -1.0, 0.0, s);
        if (n <= NT && NB == 32)
          hipLaunchKernelGGL((k_inv_panel_reg<R, 32, 1>), dim3(batch), dim3(NT), 0, s, n, c0, nbp,
                             c0 - s0, s0, sw, A, sA, w.Z, n > sw ? w.Zh : nullptr, w.Rb, w.used,
                             w.piv, active, status);
        else if (n <= NT)
          hipLaunchKernelGGL((k_inv_panel_reg<R, 16, 1>), dim3(batch), dim3(NT), 0, s, n, c0, nbp,
                             c0 - s0, s0, sw, A, sA, w.Z, n > sw ? w.Zh : nullptr, w.Rb, w.used,
                             w.piv, active, status);
        else
          hipLaunchKernelGGL((k_inv_panel_reg<R, 16, 2>), dim3(batch), dim3(NT), 0, s, n, c0, nbp,
                             c0 - s0, s0, sw, A, sA, w.Z, n > sw ? w.Zh : nullptr, w.Rb, w.used,
                             w.piv, active, status);
      }
      HN_CHECK(hipGetLastError());
    }
    if (n > sw) {  // deferred rank-sw update of every column outside the super-panel
      GemmArgs<R> g;
      g.m = n; g.n = n; g.k = sw;
      g.A = w.Zh; g.sA = (long long)n * kInvSuper; g.lda = n; g.opA = OP_N;
      g.B = w.Rb; g.sB = (long long)kInvSuper * n; g.ldb = kInvSuper; g.opB = OP_N;
      g.C = A; g.sC = sA; g.ldc = n;
      g.alpha = {1.0, 0.0}; g.beta = {1.0, 0.0};
      g.active = active; g.skip_c0 = s0; g.skip_c1 = s0 + sw;
      g.batch = batch; g.tag = PT_GEMM_INV; g.prof_items = w.prof_items;
      if constexpr (sizeof(R) == 8) gemm_z(g, s);
      else gemm_c(g, s);
    }
  }
  hipLaunchKernelGGL(k_inv_unperm<R>, dim3((n + 7) / 8, batch), dim3(NT), 2 * n * sizeof(int), s,
                     n, A, sA, Ainv, sAinv, w.piv, active);
  HN_CHECK(hipGetLastError());
}

}  // namespace

int inv_lds_max_n(int real_bytes) { return real_bytes == 8 ? lds_max_n<double>() : lds_max_n<float>(); }

size_t inv_work_bytes(int n, int batch, int real_bytes) {
  size_t c = 2 * (size_t)real_bytes;
  auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
  return al((size_t)n * kInvPanel * c * batch) + 2 * al((size_t)n * kInvSuper * c * batch) +
         al((size_t)n * batch * sizeof(int)) * 2;
}

template <typename R>
InvWork<R> inv_work_carve(void *buf, int n, int batch) {
  typedef typename InvWork<R>::CT C;
  auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
  InvWork<R> w;
  char *p = (char *)buf;
  w.Z = (C *)p; p += al((size_t)n * kInvPanel * sizeof(C) * batch);
  w.Zh = (C *)p; p += al((size_t)n * kInvSuper * sizeof(C) * batch);
  w.Rb = (C *)p; p += al((size_t)n * kInvSuper * sizeof(C) * batch);
  w.used = (int *)p; p += al((size_t)n * batch * sizeof(int));
  w.piv = (int *)p;
  w.n = n;
  w.batch = batch;
  return w;
}
template InvWork<double> inv_work_carve<double>(void *, int, int);
template InvWork<float> inv_work_carve<float>(void *, int, int);

void inv_z(int n, double2 *A, long long sA, double2 *Ainv, long long sAinv, int batch,
           const int *active, int *status, const InvWork<double> &w, hipStream_t s) {
  inv_impl<double>(n, A, sA, Ainv, sAinv, batch, active, status, w, s);
}
void inv_c(int n, float2 *A, long long sA, float2 *Ainv, long long sAinv, int batch,
           const int *active, int *status, const InvWork<float> &w, hipStream_t s) {
  inv_impl<float>(n, A, sA, Ainv, sAinv, batch, active, status, w, s);
}

}  // namespace hn
