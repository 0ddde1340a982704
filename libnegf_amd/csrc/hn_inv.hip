// hn_inv.hip -- batched complex block inverse with partial pivoting.
//
// Replaces zinv (ZGETRF + ZGETRS against I, inversions.F90:807-839) and
// cu_Zinverse (gpuroutines.cu:449-523) for whole batches of blocks.
//
// Small blocks (n^2 complex fit in LDS): one workgroup per block runs an
// in-place Gauss-Jordan elimination with partial pivoting entirely in LDS,
// followed by the column interchanges that undo the row pivoting.
//
// Large blocks: blocked in-place Gauss-Jordan with *lazy* pivoting.  Rows are
// never moved; the pivot row of each logical column is recorded instead.
// Per panel of nb columns:
//   panel kernel (one workgroup per item, panel in LDS): partial-pivoting LU
//     of the panel restricted to not-yet-pivoted rows picks nb pivot rows;
//     P = inverse of the pivot block; writes Z = [-B(r,K) P ; P - I] (n x nb),
//     a copy R of the pivot rows, and the new panel columns.
//   update: B(:, J) += Z R(:, J) for all J outside the panel -- a rank-nb
//     GEMM on the matrix cores (hn_gemm with the panel columns masked).
// After the last panel, Ainv(i, c) = B(pi(i), pi^-1(c)) with pi(c) the pivot
// row of logical column c.
#include "hn_gemm.h"
#include "hn_inv.h"

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

template <typename R, int NB>
__global__ __launch_bounds__(NT) void k_inv_panel(int n, int c0, int nbp,
                                                  typename CT_<R>::t *A, long long sA,
                                                  typename CT_<R>::t *Z,
                                                  typename CT_<R>::t *Rb, int *used_g,
                                                  int *piv_g, const int *active, int *status) {
  typedef typename CT_<R>::t C;
  extern __shared__ __align__(16) unsigned char smem[];
  const int b = blockIdx.x;
  if (active && active[b] == 0) return;
  C *P = (C *)smem;               // n x NB
  int *used = (int *)(P + (size_t)n * NB);
  int *pos = used + n;
  __shared__ C Q[NB * NB];
  __shared__ C qrow[NB], qcol[NB];
  __shared__ int piv[NB], qperm[NB];
  __shared__ R redv[NT / 64];
  __shared__ int redi[NT / 64];
  __shared__ int s_sing, s_p;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  C *Ab = A + (long long)b * sA;
  C *Zb = Z + (long long)b * n * NB;
  C *Rbb = Rb + (long long)b * NB * n;
  int *ug = used_g + (long long)b * n;
  for (int j = 0; j < nbp; ++j)
    for (int r = tid; r < n; r += NT) P[r + (size_t)j * n] = Ab[r + (size_t)(c0 + j) * n];
  for (int r = tid; r < n; r += NT) {
    used[r] = ug[r];
    pos[r] = -1;
  }
  if (tid == 0) s_sing = 0;
  __syncthreads();
  // ---- pivot selection: LU with partial pivoting over unused rows ----
  for (int j = 0; j < nbp; ++j) {
    R best = R(-1);
    int bi = -1;
    for (int r = tid; r < n; r += NT)
      if (!used[r]) {
        R v = cabs1(P[r + (size_t)j * n]);
        if (v > best || (v == best && (bi < 0 || r < bi))) { best = v; bi = r; }
      }
    wave_argmax(best, bi);
    if (lane == 0) { redv[wave] = best; redi[wave] = bi; }
    __syncthreads();
    if (tid == 0) {
      R bv = redv[0];
      int bx = redi[0];
      for (int w = 1; w < NT / 64; ++w)
        if (redv[w] > bv || (redv[w] == bv && redi[w] >= 0 && (bx < 0 || redi[w] < bx))) {
          bv = redv[w];
          bx = redi[w];
        }
      if (bx < 0) {  // every candidate is NaN: take the first unused row
        s_sing = 1;
        for (int r = 0; r < n; ++r)
          if (!used[r]) { bx = r; break; }
        if (bx < 0) bx = 0;  // unreachable: fewer than n rows pivoted so far
      }
      if (!(bv > R(0))) s_sing = 1;
      piv[j] = bx;
      used[bx] = 1;
      pos[bx] = j;
      s_p = bx;
    }
    __syncthreads();
    const int p = s_p;
    const C pinv = crecip(P[p + (size_t)j * n]);
    for (int r = tid; r < n; r += NT)
      if (!used[r]) {
        C l = cmul(P[r + (size_t)j * n], pinv);
        for (int jj = j + 1; jj < nbp; ++jj)
          P[r + (size_t)jj * n] = csub(P[r + (size_t)jj * n], cmul(l, P[p + (size_t)jj * n]));
      }
    __syncthreads();
  }
  // ---- pivot block Q = B(piv, K) (original values) and its inverse ----
  for (int e = tid; e < nbp * nbp; e += NT) {
    int jr = e % nbp, l = e / nbp;
    Q[jr + l * NB] = Ab[piv[jr] + (size_t)(c0 + l) * n];
  }
  __syncthreads();
  for (int k = 0; k < nbp; ++k) {
    if (tid == 0) {
      int p = k;
      R best = cabs1(Q[k + k * NB]);
      for (int i = k + 1; i < nbp; ++i) {
        R v = cabs1(Q[i + k * NB]);
        if (v > best) { best = v; p = i; }
      }
      if (!(best > R(0))) s_sing = 1;
      qperm[k] = p;
      if (p != k)
        for (int j = 0; j < nbp; ++j) {
          C t = Q[k + j * NB];
          Q[k + j * NB] = Q[p + j * NB];
          Q[p + j * NB] = t;
        }
    }
    __syncthreads();
    const C pinv = crecip(Q[k + k * NB]);
    if (tid < nbp) {
      qrow[tid] = (tid == k) ? pinv : cmul(Q[k + tid * NB], pinv);
      qcol[tid] = (tid == k) ? C{0, 0} : Q[tid + k * NB];
    }
    __syncthreads();
    if (tid < nbp * nbp) {
      int i = tid % nbp, j = tid / nbp;
      C v = Q[i + j * NB];
      if (i == k) v = qrow[j];
      else if (j == k) v = cneg(cmul(qcol[i], pinv));
      else v = csub(v, cmul(qcol[i], qrow[j]));
      Q[i + j * NB] = v;
    }
    __syncthreads();
  }
  for (int k = nbp - 1; k >= 0; --k) {
    const int p = qperm[k];
    if (p != k && tid < nbp) {
      C t = Q[tid + k * NB];
      Q[tid + k * NB] = Q[tid + p * NB];
      Q[tid + p * NB] = t;
    }
    __syncthreads();
  }
  // ---- copy of the pivot rows (all columns; the panel columns are unused) ----
  for (int e = tid; e < nbp * n; e += NT) {
    int j = e % nbp, c = e / nbp;
    Rbb[j + (size_t)c * NB] = Ab[piv[j] + (size_t)c * n];
  }
  __syncthreads();
  // ---- Z and the new panel columns ----
  for (int r = tid; r < n; r += NT) {
    const int j = pos[r];
    if (j >= 0) {
      for (int l = 0; l < nbp; ++l) {
        C v = Q[j + l * NB];
        Ab[r + (size_t)(c0 + l) * n] = v;
        if (l == j) v.x -= R(1);
        Zb[r + (size_t)l * n] = v;
      }
    } else {
      C row[NB];
#pragma unroll
      for (int m = 0; m < NB; ++m) row[m] = (m < nbp) ? Ab[r + (size_t)(c0 + m) * n] : C{0, 0};
      for (int l = 0; l < nbp; ++l) {
        C w = {0, 0};
#pragma unroll
        for (int m = 0; m < NB; ++m)
          if (m < nbp) {
            C q = Q[m + l * NB];
            w.x += row[m].x * q.x - row[m].y * q.y;
            w.y += row[m].x * q.y + row[m].y * q.x;
          }
        C nw = cneg(w);
        Ab[r + (size_t)(c0 + l) * n] = nw;
        Zb[r + (size_t)l * n] = nw;
      }
    }
  }
  for (int r = tid; r < n; r += NT) ug[r] = used[r];
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
    hipLaunchKernelGGL(k_inv_lds<R>, dim3(batch), dim3(NT), lds, s, n, A, sA, Ainv, sAinv,
                       active, status);
    HN_CHECK(hipGetLastError());
    return;
  }
  constexpr int NB = kInvPanel;
  typedef typename CT_<R>::t C;
  size_t lds = (size_t)n * NB * sizeof(C) + 2 * (size_t)n * sizeof(int);
  if (lds > (size_t)kLdsBudget) throw std::runtime_error("inverse: block order too large");
  if (sA < (long long)n * n || sAinv < (long long)n * n)
    throw std::runtime_error("inverse: item stride smaller than n*n");
  if (!w.Z || w.n < n || w.batch < batch) throw std::runtime_error("inverse: workspace too small");
  long long nb_tot = (long long)n * batch;
  hipLaunchKernelGGL(k_inv_reset<R>, dim3((unsigned)((nb_tot + NT - 1) / NT)), dim3(NT), 0, s,
                     n, batch, w.used, active);
  allow_lds(k_inv_panel<R, NB>, lds);
  for (int c0 = 0; c0 < n; c0 += NB) {
    int nbp = std::min(NB, n - c0);
    hipLaunchKernelGGL((k_inv_panel<R, NB>), dim3(batch), dim3(NT), lds, s, n, c0, nbp, A, sA,
                       w.Z, w.Rb, w.used, w.piv, active, status);
    HN_CHECK(hipGetLastError());
    if (n > nbp) {
      GemmArgs<R> g;
      g.m = n; g.n = n; g.k = nbp;
      g.A = w.Z; g.sA = (long long)n * NB; g.lda = n; g.opA = OP_N;
      g.B = w.Rb; g.sB = (long long)NB * n; g.ldb = NB; g.opB = OP_N;
      g.C = A; g.sC = sA; g.ldc = n;
      g.alpha = {1.0, 0.0}; g.beta = {1.0, 0.0};
      g.active = active; g.skip_c0 = c0; g.skip_c1 = c0 + nbp;
      g.batch = batch;
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
  size_t z = (size_t)n * kInvPanel * c * batch;
  auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
  return al(z) * 2 + al((size_t)n * batch * sizeof(int)) * 2;
}

template <typename R>
InvWork<R> inv_work_carve(void *buf, int n, int batch) {
  typedef typename InvWork<R>::CT C;
  auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
  InvWork<R> w;
  char *p = (char *)buf;
  size_t z = al((size_t)n * kInvPanel * sizeof(C) * batch);
  w.Z = (C *)p; p += z;
  w.Rb = (C *)p; p += z;
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
