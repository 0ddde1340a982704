// hn_inv.hip -- batched complex block inverse with partial pivoting.
//
// Replaces zinv (ZGETRF + ZGETRS against I, inversions.F90:807-839) and
// cu_Zinverse (gpuroutines.cu:449-523) for whole batches of blocks.
//
// Small blocks (n^2 complex fit in LDS): one workgroup per block runs an
// in-place Gauss-Jordan elimination with partial pivoting entirely in LDS,
// followed by the column interchanges that undo the row pivoting.
//
// Large blocks (up to n = 1024): in-place Gauss-Jordan with lazy pivoting on
// the row-major view M = A^T, so that rows of M are contiguous columns of A
// (inverting M and reading the result back column-major gives A^-1; the row
// pivoting of M is column pivoting of A).  Rows are never moved; the pivot
// row of each logical column is recorded.  Per super-panel S of up to 64
// columns, one launch (one workgroup of 512 threads per item, the n x |S|
// panel in registers) runs the |S| elimination steps on S and writes the
// accumulated transform; the other columns of M then take ONE deferred
// rank-|S| update on the matrix cores.  After the last super-panel,
// M^-1(i, c) = B(pi(i), pi^-1(c)) with pi(c) the pivot row of column c.
#include <utility>

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
constexpr int NTG = 512;  // threads of the Gauss-Jordan kernel: two halves of 256 rows

// Encoding of a chosen pivot row r (r < 1024) with its FP64 |.|_1 (the row
// itself comes from the exact 64-bit search): the high word of the magnitude
// (exponent and 9 mantissa bits), a flag bit and 1023 - r; the row is read
// back from the low bits and a magnitude below 2^-1011 (the high word below
// 0x800) flags a zero pivot.
__device__ inline unsigned pivot_key(double v, int r) {
  const unsigned hi = min((unsigned)((unsigned long long)__double_as_longlong(v) >> 32), 0x7FF00000u);
  return (hi & ~0x7FFu) | 0x400u | (unsigned)(1023 - r);
}
__device__ inline unsigned pivot_key(float v, int r) {
  return pivot_key((double)v, r);
}
// 1/b for a pivot: b scaled by a power of two (exact) so that |b'|^2 lies in
// [1/4, 2), the hardware reciprocal of |b'|^2 refined by two Newton-Raphson
// steps (r <- r + r (1 - d r)), and conj(b') r scaled back; 0 for b == 0.
__device__ inline double2 crecip_fast(double2 b) {
  const double mx = fmax(fabs(b.x), fabs(b.y));
  if (mx == 0.0) return double2{0, 0};
  const int e = __builtin_amdgcn_frexp_exp(mx);
  const double x = __builtin_amdgcn_ldexp(b.x, -e), y = __builtin_amdgcn_ldexp(b.y, -e);
  const double d = fma(x, x, y * y);
  double r = __builtin_amdgcn_rcp(d);
  r = fma(r, fma(-d, r, 1.0), r);
  r = fma(r, fma(-d, r, 1.0), r);
  return double2{__builtin_amdgcn_ldexp(x * r, -e), __builtin_amdgcn_ldexp(-y * r, -e)};
}
__device__ inline float2 crecip_fast(float2 b) { return crecip(b); }

// value held by lane `lane` (uniform) of the wave
__device__ inline double lane_read(double v, int lane) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
  return __hiloint2double(hi, lo);
}
__device__ inline float lane_read(float v, int lane) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

// 64-bit key helpers for the exact pivot search: DPP butterflies inside each
// row of 16 lanes on both halves, then the four row results read as scalars.
template <int CTRL> __device__ inline unsigned long long dpp64(unsigned long long k) {
  const unsigned hi = __builtin_amdgcn_update_dpp(0, (int)(k >> 32), CTRL, 0xF, 0xF, false);
  const unsigned lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)k, CTRL, 0xF, 0xF, false);
  return ((unsigned long long)hi << 32) | lo;
}
__device__ inline unsigned long long rl64(unsigned long long k, int lane) {
  const unsigned hi = __builtin_amdgcn_readlane((int)(k >> 32), lane);
  const unsigned lo = __builtin_amdgcn_readlane((int)(unsigned)k, lane);
  return ((unsigned long long)hi << 32) | lo;
}
__device__ inline unsigned long long wave_max64(unsigned long long k) {
  k = max(k, dpp64<0xB1>(k));   // quad_perm [1,0,3,2]
  k = max(k, dpp64<0x4E>(k));   // quad_perm [2,3,0,1]
  k = max(k, dpp64<0x124>(k));  // row_ror:4
  k = max(k, dpp64<0x128>(k));  // row_ror:8
  return max(max(rl64(k, 0), rl64(k, 16)), max(rl64(k, 32), rl64(k, 48)));
}
template <int CTRL> __device__ inline unsigned dpp_min(unsigned k) {
  return min(k, (unsigned)__builtin_amdgcn_update_dpp(-1, (int)k, CTRL, 0xF, 0xF, false));
}
__device__ inline unsigned wave_min32(unsigned k) {
  k = dpp_min<0xB1>(k);
  k = dpp_min<0x4E>(k);
  k = dpp_min<0x124>(k);
  k = dpp_min<0x128>(k);
  const unsigned a = __builtin_amdgcn_readlane((int)k, 0), b = __builtin_amdgcn_readlane((int)k, 16);
  const unsigned c = __builtin_amdgcn_readlane((int)k, 32), d = __builtin_amdgcn_readlane((int)k, 48);
  return min(min(a, b), min(c, d));
}

// ---------------------------------------------------------------------------
// LU inverse of blocks up to 64 (complex double), one workgroup per block
// ---------------------------------------------------------------------------
// The reference's algorithm: ZGETRF (partial pivoting; the pivot is the first
// largest |Re| + |Im| of the column in the current row order, as IZAMAX
// picks it) followed by ZGETRS against the identity (inversions.F90:807-839).
//
// G x G threads, thread (cb, rb) = threadIdx cb*G + rb holding the TB x TB
// sub-block rows rb*TB.., columns cb*TB.. in registers.  The G threads of one
// thread column are G consecutive lanes of one wave.  Rows are never moved:
// pos is a row's current position in LAPACK's row order, and a pivot step
// exchanges two positions.  One register block holds both factors in place:
// columns right of the current step hold A, columns at or left of it hold
// Y = L^-1 P (the forward substitution of ZGETRS against P I; column j of Y
// belongs to the row pivoted at step j, so the pivoted rows' entries of Y
// right of their own step are zero).
//
// Step k: the pivot p (row, position, reciprocal) and column k after step
// k-1 are in LDS.  The thread column owning column k turns that register
// column into Y's column k (1 on row p, 0 elsewhere).  Every thread takes row
// p for its columns from the lane of its own thread column that holds it
// (ds_bpermute inside the wave) and applies X(r, c) -= m_r X(p, c),
// m_r = A(r, k)/A(p, k), to its unpivoted rows -- ZGETRF's elimination on
// A's columns and ZGETRS's forward substitution on Y's, in one pass with no
// branches.  Then the thread column owning column k+1 publishes it: the
// entries of pivoted rows (U above the diagonal) into a triangular LDS array
// by position, the others for the next multipliers, and the pivot -- the
// largest |Re| + |Im| over the unpivoted rows as an exact 64-bit key, ties to
// the smallest position, reduced by DPP inside the G lanes.  One barrier per
// step.
//
// Every register index is a compile-time constant (static_for, the k loop
// unrolled by TB, and run-time rows read through select chains), so the
// block stays in registers.
//
// Back substitution U X = Y, positions i = n-1 .. 0, without barriers: row i
// of X (the row pivoted at step i, divided by U(i, i)) from the lanes of the
// same thread column, U's column i from the triangular array; the division
// of a row's own registers is applied at the end (a row is not updated after
// its own position).  Row pos(r) of X ends in the threads of row r; column j of X is
// column perm(j) of the inverse (the row pivoted at step j).

template <typename F, int... I>
__device__ __forceinline__ void static_for_impl(F &&f, std::integer_sequence<int, I...>) {
  (f(std::integral_constant<int, I>{}), ...);
}
// f(std::integral_constant<int, i>) for i = 0 .. N-1, unrolled at compile time
template <int N, typename F> __device__ __forceinline__ void static_for(F &&f) {
  static_for_impl(f, std::make_integer_sequence<int, N>{});
}

// (key, pos) pair reduction over G consecutive lanes (G = 4, 8 or 16): the
// larger key wins, equal keys go to the smaller pos
template <int CTRL> __device__ inline void best_dpp(unsigned long long &k, int &p) {
  const unsigned hi = __builtin_amdgcn_update_dpp(0, (int)(k >> 32), CTRL, 0xF, 0xF, false);
  const unsigned lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)k, CTRL, 0xF, 0xF, false);
  const int op = __builtin_amdgcn_update_dpp(0, p, CTRL, 0xF, 0xF, false);
  const unsigned long long ok = ((unsigned long long)hi << 32) | lo;
  if (ok > k || (ok == k && op < p)) {
    k = ok;
    p = op;
  }
}
template <int G> __device__ inline void group_best(unsigned long long &k, int &p) {
  best_dpp<0xB1>(k, p);                  // quad_perm [1,0,3,2]
  best_dpp<0x4E>(k, p);                  // quad_perm [2,3,0,1]
  if (G >= 8) best_dpp<0x141>(k, p);     // row_half_mirror
  if (G >= 16) best_dpp<0x140>(k, p);    // row_mirror
}
// x - m y with the FMA order of the elimination updates
__device__ inline double2 fms_c(double2 x, double2 m, double2 y) {
  x.x = fma(-m.x, y.x, x.x);
  x.x = fma(m.y, y.y, x.x);
  x.y = fma(-m.x, y.y, x.y);
  x.y = fma(-m.y, y.x, x.y);
  return x;
}
// value `v` of lane `src` (lane index within the wave), for every lane
__device__ inline double bperm(double v, int src) {
  const int lo = __builtin_amdgcn_ds_bpermute(src << 2, __double2loint(v));
  const int hi = __builtin_amdgcn_ds_bpermute(src << 2, __double2hiint(v));
  return __hiloint2double(hi, lo);
}
__device__ inline double2 bperm(double2 v, int src) {
  return double2{bperm(v.x, src), bperm(v.y, src)};
}
// offset of U(i, c), i < c, in the triangular LDS array (by position)
__device__ inline int utri(int i, int c) { return (c * (c - 1)) / 2 + i; }

template <int TB> struct LuBlk {
  double2 v[TB][TB];
  int pos[TB];
  bool done[TB];
};

// element (r, J) of the register block for the uniform run-time row r < TB,
// as a select chain that starts from zero (so it is not folded into a
// register-indexed load, which would move the block out of registers)
template <int TB, int J>
__device__ __forceinline__ double2 pick0(const double2 (&v)[TB][TB], int r) {
  double xr = 0.0, xi = 0.0;
  static_for<TB>([&](auto I) {
    const bool hit = r == (int)I;
    xr = hit ? v[I][J].x : xr;
    xi = hit ? v[I][J].y : xi;
  });
  return double2{xr, xi};
}

// The thread column owning column col (register column JJ) publishes it:
// pivoted rows' entries into the triangular U array, unpivoted rows' into
// colbuf (the next multipliers), and the pivot's row, position and
// reciprocal.  Called by every thread; only the owner's wave works.
template <int G, int TB, int JJ>
__device__ __forceinline__ void lu_publish(int col, int tid, const LuBlk<TB> &B, double2 *Ut,
                                           double2 *colbuf, int *s_perm, int *s_pospiv,
                                           double2 *s_uinv, int *s_sing) {
  const int cb = tid / G, rb = tid % G;
  const int oc = col / TB;
  if ((tid >> 6) != (oc * G) >> 6) return;  // wave-uniform: only the owner's wave
  unsigned long long key = 0, kk[TB];
  int bp = 0x7fffffff;
  static_for<TB>([&](auto I) {
    constexpr int i = I;
    const double2 v = B.v[i][JJ];
    kk[i] = B.done[i] ? 0ull
                      : (unsigned long long)__double_as_longlong(fabs(v.x) + fabs(v.y)) + 1ull;
    if (cb == oc) {
      if (B.done[i]) Ut[utri(B.pos[i], col)] = v;
      else colbuf[rb * TB + i] = v;
    }
    if (kk[i] > key || (kk[i] == key && B.pos[i] < bp)) {
      key = kk[i];
      bp = B.pos[i];
    }
  });
  group_best<G>(key, bp);
  if (cb == oc) {
    static_for<TB>([&](auto I) {
      constexpr int i = I;
      if (!B.done[i] && kk[i] == key && B.pos[i] == bp) {
        s_perm[col] = rb * TB + i;
        s_pospiv[col & 1] = bp;
        s_uinv[col] = crecip_fast(B.v[i][JJ]);
        if (key <= 1ull) *s_sing = 1;  // the largest candidate is zero
      }
    });
  }
}

template <int G, int TB>
__global__ __launch_bounds__(G * G) void k_lu_inv_2d(int n, const double2 *__restrict__ A,
                                                     long long sA, double2 *__restrict__ Ainv,
                                                     long long sAinv,
                                                     const int *__restrict__ active,
                                                     int *__restrict__ status) {
  typedef double2 C;
  constexpr int LD = G * TB;  // rows (and columns) of the padded block
  static_assert(G * G >= 64 || G * G == 16, "thread columns inside one wave");
  const int b = blockIdx.x;
  if (active && active[b] == 0) return;
  extern __shared__ __align__(16) unsigned char lsm[];
  C *Ut = (C *)lsm;  // U(i, c), i < c (positions), at utri(i, c)
  __shared__ int s_perm[LD];
  __shared__ C s_uinv[LD];
  __shared__ int s_pospiv[2];
  __shared__ int s_sing;
  __shared__ C colbuf[2][LD];
  const int tid = threadIdx.x, cb = tid / G, rb = tid % G;
  const int gbase = (tid & 63) - rb;  // lane of this thread column's rb = 0
  const C *Ab = A + (long long)b * sA;
  LuBlk<TB> B;
  static_for<TB>([&](auto I) {
    constexpr int i = I;
    const int r = rb * TB + i;
    B.pos[i] = r;
    B.done[i] = r >= n;
    static_for<TB>([&](auto J) {
      const int c = cb * TB + J;
      B.v[i][J] = (r < n && c < n) ? Ab[r + (size_t)c * n] : C{0, 0};
    });
  });
  if (tid == 0) s_sing = 0;
  lu_publish<G, TB, 0>(0, tid, B, Ut, colbuf[0], s_perm, s_pospiv, s_uinv, &s_sing);
  __syncthreads();

  for (int k0 = 0; k0 < n; k0 += TB) {
    static_for<TB>([&](auto KK) {
      constexpr int kk = KK;
      const int k = k0 + kk;
      if (k >= n) return;  // uniform
      const int p = s_perm[k], pos_p = s_pospiv[k & 1];
      const C pinv = s_uinv[k];
      const int prb = p / TB, pi = p - prb * TB;
      const bool own_k = cb == k0 / TB;
      C m[TB];
      static_for<TB>([&](auto I) {
        constexpr int i = I;
        const int r = rb * TB + i;
        if (r == p) {
          B.pos[i] = k;
          B.done[i] = true;
        } else if (B.pos[i] == k) {
          B.pos[i] = pos_p;
        }
        m[i] = B.done[i] ? C{0, 0} : cmul(colbuf[k & 1][r], pinv);
        // register column k becomes Y's column k: 1 on the pivot row
        if (own_k) B.v[i][kk] = C{r == p ? 1.0 : 0.0, 0.0};
      });
      // row p for this thread's columns, from the lane of this thread column
      // that holds it (same wave)
      const int ps = __builtin_amdgcn_readfirstlane(pi);
      static_for<TB>([&](auto J) {
        const C u = bperm(pick0<TB, J>(B.v, ps), gbase + prb);
        static_for<TB>([&](auto I) { B.v[I][J] = fms_c(B.v[I][J], m[I], u); });
      });
      if (k + 1 < n)
        lu_publish<G, TB, (kk + 1) % TB>(k + 1, tid, B, Ut, colbuf[(k + 1) & 1], s_perm,
                                         s_pospiv, s_uinv, &s_sing);
      __syncthreads();
    });
  }

  // back substitution U X = Y, position i = n-1 .. 0, each wave on its own
  // columns: row i of X (row q of Y / U(i, i), left unscaled in the registers
  // of row q) from the lane of this thread column that holds row q
  for (int i = n - 1; i >= 0; --i) {
    const int q = s_perm[i];
    const C uinv = s_uinv[i];
    const int qrb = q / TB, qs = __builtin_amdgcn_readfirstlane(q - qrb * TB);
    C coef[TB];
    static_for<TB>([&](auto T) {
      coef[T] = (B.pos[T] < i) ? Ut[utri(B.pos[T], i)] : C{0, 0};
    });
    static_for<TB>([&](auto J) {
      const C xq = bperm(cmul(pick0<TB, J>(B.v, qs), uinv), gbase + qrb);
      // coef = 0 on row q and below position i
      static_for<TB>([&](auto T) { B.v[T][J] = fms_c(B.v[T][J], coef[T], xq); });
    });
  }
  // rows of X: the scaling by 1 / U(pos, pos) left out above
  static_for<TB>([&](auto I) {
    if (rb * TB + I < n) {
      const C ui = s_uinv[B.pos[I]];
      static_for<TB>([&](auto J) { B.v[I][J] = cmul(B.v[I][J], ui); });
    }
  });
  // column j of X = U^-1 L^-1 is column perm(j) of the inverse
  C *Ob = Ainv + (long long)b * sAinv;
  static_for<TB>([&](auto J) {
    const int c = cb * TB + J;
    if (c < n) {
      const int pc = s_perm[c];
      static_for<TB>([&](auto I) {
        if (rb * TB + I < n) Ob[B.pos[I] + (size_t)pc * n] = B.v[I][J];
      });
    }
  });
  if (tid == 0 && s_sing && status) atomicOr(&status[b], (int)ST_SINGULAR);
}


// "row already pivoted" flags of a lane's RT register rows: an array, or at
// 16 rows per lane the bits of one register (the panel then holds 128)
template <int RT, bool BITS = (RT >= 16)> struct RowFlags {
  int f[RT];
  __device__ __forceinline__ int get(int i) const { return f[i]; }
  __device__ __forceinline__ void set(int i, int v) { f[i] = v; }
};
template <int RT> struct RowFlags<RT, true> {
  static_assert(RT <= 32, "one 32-bit mask");
  unsigned m = 0;
  __device__ __forceinline__ int get(int i) const { return (int)((m >> i) & 1u); }
  __device__ __forceinline__ void set(int i, int v) { m = v ? (m | (1u << i)) : (m & ~(1u << i)); }
};

// rows of the LDS staging buffer of k_inv_gj (a chunk of the n x 8CT panel,
// rows padded by one element); at 16 rows per lane the multiplier buffers
// (2 x 64 RT elements) take their share of the budget
template <typename R, int RT, int CT>
constexpr int gj_stage_rows() {
  constexpr int rowb = (8 * CT + 1) * 2 * (int)sizeof(R);
  constexpr int budget = 140 * 1024 - (RT >= 16 ? 2 * 64 * RT * 2 * (int)sizeof(R) : 0);
  constexpr int cr = (budget / rowb) & ~63;
  return cr < 64 * RT ? cr : 64 * RT;
}
template <typename R, int RT, int CT>
constexpr size_t gj_stage_bytes() {
  return (size_t)gj_stage_rows<R, RT, CT>() * (8 * CT + 1) * 2 * sizeof(R);
}

// Gauss-Jordan elimination over one super-panel S = columns [s0, s0+sw),
// sw <= 8 CT, of the row-major view M = A^T of an item (row r of M is column r
// of A, so rows of M are contiguous in memory).
//
// Wave w (of 8) holds columns s0 + w*CT .. +CT, lane l the rows l + 64 i
// (i < RT), in registers (loaded and stored through an LDS staging buffer so
// that global accesses run along rows).  Every step writes each updated value
// one register lower (register 0 moves to the last), so the wave owning the
// current pivot column always finds it in register 0 and no register is
// indexed at run time; CT steps restore the order.
//
// Step j, pivot column owned by wave wj = j / CT, one barrier:
//   every thread reads the pivot key (row p, the largest |.|_1 over the
//   unpivoted rows) and 1/B(p, j); each wave reads its CT entries of row p
//   from the lane holding them (1 in place of column j); every row applies
//   B(r, :) -= (B(r, j)/B(p, j)) B(p, :), the multipliers of the other waves'
//   rows coming from LDS, and the pivot row becomes B(p, :)/B(p, j) (1/B(p, j)
//   in column j).  The wave owning the next pivot column updates that column
//   first and, before its other columns, publishes the column (the next
//   multipliers) and its pivot candidate with the reciprocal, so the search
//   overlaps the update.  LDS buffers alternate between steps.
//
// With Ẑ = B(:, S) - E (E: ones at (pivot row, column)) the columns J outside
// S then take B(:, J) += Ẑ B_old(piv, J).  In A's column-major layout that is
// A(J', :) += Rb Ẑ^T with Rb(c, k) = A(c, piv_k) -- contiguous columns of A --
// and J' the rows of A outside S.  Ẑ^T is the rows S of A less E^T, and
// -Rb E^T only cancels A(J', piv_k): the kernel moves those entries into Rb
// (ld n; rows of S are not written) and leaves zeros, so that the deferred
// update on the matrix cores is the plain product A(J', :) += Rb A(S, :).
template <typename R, int RT, int CT>
__global__ __launch_bounds__(NTG) void k_inv_gj(int n, int s0, int sw,
                                                typename CT_<R>::t *__restrict__ A, long long sA,
                                                typename CT_<R>::t *__restrict__ Rb,
                                                int *__restrict__ used_g, int *__restrict__ piv_g,
                                                const int *__restrict__ active,
                                                int *__restrict__ status) {
  typedef typename CT_<R>::t C;
  static_assert(CT >= 2, "the look-ahead needs two columns per wave");
  constexpr int SWC = 8 * CT, LDST = SWC + 1;
  constexpr int CR = gj_stage_rows<R, RT, CT>();
  constexpr int PER = CR * SWC / NTG;  // staged elements per thread and chunk
  const int b = blockIdx.x;
  if (active && active[b] == 0) return;
  extern __shared__ __align__(16) unsigned char gsm[];
  C *stage = (C *)gsm;
  __shared__ C mcol[2][64 * RT];
  __shared__ unsigned redk[2];
  __shared__ C redinv[2];
  __shared__ int spiv[SWC];
  __shared__ int s_sing;
  const int tid = threadIdx.x, w = tid >> 6, l = tid & 63;
  C *Xb = A + (long long)b * sA;
  int *ug = used_g + (long long)b * n;

  C P[RT][CT];
  RowFlags<RT> used;
#pragma unroll
  for (int i = 0; i < RT; ++i) {
    const int r = l + 64 * i;
    used.set(i, r >= n ? 1 : (s0 == 0 ? 0 : ug[r]));
#pragma unroll
    for (int c = 0; c < CT; ++c) P[i][c] = C{0, 0};
  }
  if (tid == 0) s_sing = 0;
  // ---- load the panel in chunks of CR rows, NPF chunks' global loads
  // issued together (one at 16 rows per lane, where the panel itself holds
  // 128 registers), coalesced along rows of M ----
  static_assert(CR % 64 == 0, "a lane's rows fall in one chunk each");
  constexpr int NCH = (64 * RT + CR - 1) / CR;
  constexpr int NPF = RT >= 16 ? 1 : 2;
#pragma unroll
  for (int h0 = 0; h0 < NCH; h0 += NPF) {
  C ld[NPF][PER];
#pragma unroll
  for (int h = 0; h < NPF; ++h)
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int e = tid + NTG * k;
      const int row = (h0 + h) * CR + e / SWC, col = e % SWC;
      ld[h][k] = (h0 + h < NCH && row < n && col < sw) ? Xb[(size_t)row * n + s0 + col] : C{0, 0};
    }
#pragma unroll
  for (int h = 0; h < NPF; ++h) {
    const int cr0 = (h0 + h) * CR;
    if (h0 + h < NCH && cr0 < n) {  // uniform
#pragma unroll
      for (int k = 0; k < PER; ++k) {
        const int e = tid + NTG * k;
        stage[(e / SWC) * LDST + e % SWC] = ld[h][k];
      }
      __syncthreads();
#pragma unroll
      for (int i = 0; i < RT; ++i) {
        const int r = l + 64 * i;
        if (r >= cr0 && r < cr0 + CR) {
          const C *src = stage + (r - cr0) * LDST + w * CT;
#pragma unroll
          for (int c = 0; c < CT; ++c) P[i][c] = src[c];
        }
      }
      __syncthreads();
    }
  }
  }

  // the owning wave's column (register 0) -> multipliers in mcol[bb] and the
  // pivot candidate (largest key over the wave's unpivoted rows) with its
  // reciprocal in redk[bb] / redinv[bb]
  // exact search: the largest |Re| + |Im| (its FP64 bit pattern as a 64-bit
  // key), ties to the smallest row; the winner publishes its row as a
  // pivot_key (decoded by every wave) and its reciprocal
  auto candidate = [&](int bb) {
    unsigned long long kb = 0;
    int rb = 0x7fffffff;
    double vb = 0.0;
    C cv = C{0, 0};
#pragma unroll
    for (int i = 0; i < RT; ++i) {
      const int r = l + 64 * i;
      if (r < n) mcol[bb][r] = P[i][0];
      const double v = (double)cabs1(P[i][0]);
      const unsigned long long key =
          used.get(i) ? 0ull : (unsigned long long)__double_as_longlong(v) + 1ull;
      if (key > kb) { kb = key; rb = r; vb = v; cv = P[i][0]; }  // rows grow with i
    }
    const unsigned long long km = wave_max64(kb);
    const unsigned rmin = wave_min32(kb == km ? (unsigned)rb : 0xFFFFFFFFu);
    if (km != 0 && kb == km && (unsigned)rb == rmin) {
      redk[bb] = pivot_key(vb, rb);
      redinv[bb] = crecip_fast(cv);
    }
    if (km == 0 && l == 0) redk[bb] = 0;
  };
  if (w == 0) candidate(0);
  __syncthreads();

  for (int j = 0; j < sw; ++j) {
    const int bf = j & 1, wj = j / CT;
    // ---- pivot row p and 1/B(p, j); each wave reads its columns of row p
    // from the lane holding it ----
    const unsigned kx = redk[bf];
    const int p = kx ? 1023 - (int)(kx & 1023) : -1;
    if (tid == 0 && (kx >> 11) == 0) s_sing = 1;  // zero (or NaN) pivot
    const C pinv = kx ? redinv[bf] : C{0, 0};
    const int lp = p & 63, ip = p >> 6;  // p < 0 cannot happen: j < n unpivoted rows remain
    C y[CT];
#pragma unroll
    for (int c = 0; c < CT; ++c) y[c] = C{0, 0};
#pragma unroll
    for (int i = 0; i < RT; ++i)
      if (p >= 0 && ip == i) {
#pragma unroll
        for (int c = 0; c < CT; ++c) y[c] = C{lane_read(P[i][c].x, lp), lane_read(P[i][c].y, lp)};
        if (l == lp) used.set(i, 1);
      }
    if (w == wj) y[0] = C{R(1), R(0)};  // column j itself
    if (w == wj && l == 0) spiv[j] = p < 0 ? 0 : p;
    // ---- B(r, :) -= l_r B(p, :), l_r = B(r, j) / B(p, j); old register k + 1
    // -> register k (the owning wave's column j, taken as 0, enters the last
    // register against the 1 in y[0]) ----
    if constexpr (RT >= 16) {
      // 16 rows per lane: row by row (multiplier, then the row's CT registers),
      // so no per-row arrays are held; the next column is published after the
      // whole update.  The same operations on every element as below.
#pragma unroll
      for (int i = 0; i < RT; ++i) {
        const int r = l + 64 * i;
        const C m = (w == wj) ? P[i][0] : (r < n ? mcol[bf][r] : C{0, 0});
        C lmi = cmul(m, pinv);
        C x0i = (w == wj) ? C{0, 0} : P[i][0];
        if (r == p) {
          lmi = C{-pinv.x, -pinv.y};
          x0i = C{0, 0};
#pragma unroll
          for (int c = 1; c < CT; ++c) P[i][c] = C{0, 0};
        }
        const R nr = -lmi.x, ni = -lmi.y;
#pragma unroll
        for (int k = 0; k < CT; ++k) {
          const C yk = y[(k + 1) % CT];
          C x = k + 1 < CT ? P[i][k + 1] : x0i;
          x.x = fma(nr, yk.x, x.x);
          x.x = fma(-ni, yk.y, x.x);
          x.y = fma(nr, yk.y, x.y);
          x.y = fma(ni, yk.x, x.y);
          P[i][k] = x;
        }
      }
      if (j + 1 < sw && w == (j + 1) / CT) candidate(bf ^ 1);
      __syncthreads();
      continue;
    }
    C lm[RT], x0[RT];
#pragma unroll
    for (int i = 0; i < RT; ++i) {
      const int r = l + 64 * i;
      const C m = (w == wj) ? P[i][0] : (r < n ? mcol[bf][r] : C{0, 0});
      lm[i] = cmul(m, pinv);
      x0[i] = (w == wj) ? C{0, 0} : P[i][0];
      if (r == p) {  // the pivot row becomes B(p, :) / B(p, j)
        lm[i] = C{-pinv.x, -pinv.y};
        x0[i] = C{0, 0};
#pragma unroll
        for (int c = 1; c < CT; ++c) P[i][c] = C{0, 0};
      }
    }
    auto upd = [&](int k) {
      const C yk = y[(k + 1) % CT];
#pragma unroll
      for (int i = 0; i < RT; ++i) {
        C x = k + 1 < CT ? P[i][k + 1] : x0[i];
        const R nr = -lm[i].x, ni = -lm[i].y;
        x.x = fma(nr, yk.x, x.x);
        x.x = fma(-ni, yk.y, x.x);
        x.y = fma(nr, yk.y, x.y);
        x.y = fma(ni, yk.x, x.y);
        P[i][k] = x;
      }
    };
    upd(0);  // the next pivot column of its wave lands in register 0
    if (j + 1 < sw && w == (j + 1) / CT) candidate(bf ^ 1);
#pragma unroll
    for (int k = 1; k < CT; ++k) upd(k);
    __syncthreads();
  }
  // restore the register order after a narrow super-panel
  for (int j = sw; j % CT != 0; ++j) {
#pragma unroll
    for (int i = 0; i < RT; ++i) {
      const C first = P[i][0];
#pragma unroll
      for (int c = 1; c < CT; ++c) P[i][c - 1] = P[i][c];
      P[i][CT - 1] = first;
    }
  }

  // ---- write the super-panel back, CR rows at a time ----
  for (int cr0 = 0; cr0 < n; cr0 += CR) {
    const int nr = min(CR, n - cr0);
    if (cr0 > 0) __syncthreads();  // the previous chunk has been read
#pragma unroll
    for (int i = 0; i < RT; ++i) {
      const int r = l + 64 * i;
      if (r >= cr0 && r < cr0 + nr) {
        C *dst = stage + (r - cr0) * LDST + w * CT;
#pragma unroll
        for (int c = 0; c < CT; ++c) dst[c] = P[i][c];
      }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int e = tid + NTG * k;
      const int row = e / SWC, col = e % SWC;
      if (row < nr && col < sw) Xb[(size_t)(cr0 + row) * n + s0 + col] = stage[row * LDST + col];
    }
  }
#pragma unroll
  for (int i = 0; i < RT; ++i) {
    const int r = l + 64 * i;
    if (w == 0 && r < n) ug[r] = used.get(i);
  }
  if (tid < sw) piv_g[(long long)b * n + s0 + tid] = spiv[tid];
  if (Rb) {  // pivot columns of A outside the rows of S, moved to Rb; element e = (c, k)
    C *Rbb = Rb + (long long)b * kInvSuper * n;
    constexpr int GU = 8;  // loads in flight per thread
    const float inv_n = 1.0f / (float)n;  // e / n below is exact: e < 2^15, n <= 1024
    const int tot = sw * n;
    for (int base = 0; base < tot; base += NTG * GU) {
      C v[GU];
      long long o[GU];
#pragma unroll
      for (int u = 0; u < GU; ++u) {
        const int e = base + tid + NTG * u;
        const int k = (int)(((float)e + 0.5f) * inv_n), c = e - k * n;
        o[u] = (e < tot && (c < s0 || c >= s0 + sw)) ? c + (long long)spiv[k] * n : -1;
        v[u] = o[u] >= 0 ? Xb[o[u]] : C{0, 0};
      }
#pragma unroll
      for (int u = 0; u < GU; ++u)
        if (o[u] >= 0) {
          Rbb[base + tid + NTG * u] = v[u];
          Xb[o[u]] = C{0, 0};
        }
    }
  }
  if (tid == 0 && s_sing && status) atomicOr(&status[b], (int)ST_SINGULAR);
}

// Ainv(c, i) = B(ip[c], pv[i]): column i of the inverse is column pv[i] of
// the working matrix with its rows taken in the order ip (pv: pivot row of
// each logical column of M = A^T, ip its inverse permutation).
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
  // output columns handled by this workgroup: blockIdx.x * 8 .. +8
  for (int ii = 0; ii < 8; ++ii) {
    const int i = blockIdx.x * 8 + ii;
    if (i >= n) break;
    const C *src = Bb + (size_t)pv[i] * n;
    for (int c = threadIdx.x; c < n; c += NT) Ob[c + (size_t)i * n] = src[ip[c]];
  }
}

template <typename R>
size_t lds_bytes_small(int n) {
  typedef typename CT_<R>::t C;
  return (size_t)n * n * sizeof(C) + 2 * (size_t)n * sizeof(C) + (size_t)n * sizeof(int);
}

// largest n for the one-workgroup LDS kernel (FP32): what fits
template <typename R>
int lds_max_n() {
  int n = 1;
  while (n + 1 <= NT && lds_bytes_small<R>(n + 1) <= (size_t)kLdsBudget) ++n;
  return n;
}
// largest FP64 order of the register LU kernel: 16 x 16 threads with 6 x 6
// blocks (250 VGPRs, the U array 73 KiB at n = 96, two workgroups per CU).
// Above it the blocked path; 2000 items of order 65 / 80 / 96 take 1.0 / 1.2 /
// 1.7 ms here against 1.7 / 2.2 / 2.6 ms on the LDS or blocked kernels.
constexpr int kLuMaxN = 96;

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
  // algorithmic work of the one-launch inverses: 8 n^3 flops and the block read
  // and written per item (items unknown for a mask without a count: -1)
  const double items = w.prof_items >= 0 ? (double)w.prof_items : (active ? -1.0 : (double)batch);
  const double sm_fl = items < 0 ? -1.0 : 8.0 * (double)n * n * n * items;
  const double sm_by = items < 0 ? 0.0 : 2.0 * sizeof(typename CT_<R>::t) * (double)n * n * items;
  if constexpr (sizeof(R) == 8) {
    if (n <= kLuMaxN) {
      ProfScope ps(PT_INV_SMALL, sm_fl, sm_by, s);
      // the triangular U array: n (n - 1) / 2 entries
      const size_t lds = std::max<size_t>(1, (size_t)n * (n - 1) / 2) * sizeof(double2);
      if (n <= 16)
        hipLaunchKernelGGL((k_lu_inv_2d<8, 2>), dim3(batch), dim3(64), lds, s, n, A, sA, Ainv,
                           sAinv, active, status);
      else if (n <= 32)
        hipLaunchKernelGGL((k_lu_inv_2d<8, 4>), dim3(batch), dim3(64), lds, s, n, A, sA, Ainv,
                           sAinv, active, status);
      else if (n <= 64)
        hipLaunchKernelGGL((k_lu_inv_2d<16, 4>), dim3(batch), dim3(256), lds, s, n, A, sA, Ainv,
                           sAinv, active, status);
      else if (n <= 80) {
        allow_lds(k_lu_inv_2d<16, 5>, lds);
        hipLaunchKernelGGL((k_lu_inv_2d<16, 5>), dim3(batch), dim3(256), lds, s, n, A, sA, Ainv,
                           sAinv, active, status);
      } else {
        allow_lds(k_lu_inv_2d<16, 6>, lds);
        hipLaunchKernelGGL((k_lu_inv_2d<16, 6>), dim3(batch), dim3(256), lds, s, n, A, sA, Ainv,
                           sAinv, active, status);
      }
      HN_CHECK(hipGetLastError());
      return;
    }
  }
  if (sizeof(R) == 4 && n <= lds_max_n<R>()) {  // FP32; FP64 above kLuMaxN is blocked
    size_t lds = lds_bytes_small<R>(n);
    allow_lds(k_inv_lds<R>, lds);
    ProfScope ps(PT_INV_SMALL, sm_fl, sm_by, s);
    hipLaunchKernelGGL(k_inv_lds<R>, dim3(batch), dim3(NT), lds, s, n, A, sA, Ainv, sAinv,
                       active, status);
    HN_CHECK(hipGetLastError());
    return;
  }
  if (n > 1024) throw std::runtime_error("inverse: block order > 1024 is not supported");
  if (sA < (long long)n * n || sAinv < (long long)n * n)
    throw std::runtime_error("inverse: item stride smaller than n*n");
  if (!w.Rb || w.n < n || w.batch < batch) throw std::runtime_error("inverse: workspace too small");
  // register tile per thread: 4 rows x 8 columns (super-panel 64) up to
  // n = 256, 8 rows x 4 columns (super-panel 32) up to n = 512, 16 rows x 2
  // columns (super-panel 16) up to n = 1024
  const int rpt = n <= 256 ? 1 : n <= 512 ? 2 : 4;
  const int spw = rpt == 1 ? 64 : rpt == 2 ? 32 : 16;
  for (int s0 = 0; s0 < n; s0 += spw) {
    const int sw = std::min(spw, n - s0);
    typename CT_<R>::t *rb = n > sw ? w.Rb : nullptr;
    {
      ProfScope ps(PT_INV_PANEL, -1.0, 0.0, s);
      if (rpt == 1) {
        allow_lds(k_inv_gj<R, 4, 8>, gj_stage_bytes<R, 4, 8>());
        hipLaunchKernelGGL((k_inv_gj<R, 4, 8>), dim3(batch), dim3(NTG),
                           (gj_stage_bytes<R, 4, 8>()), s, n, s0, sw, A, sA, rb, w.used,
                           w.piv, active, status);
      } else if (rpt == 2) {
        allow_lds(k_inv_gj<R, 8, 4>, gj_stage_bytes<R, 8, 4>());
        hipLaunchKernelGGL((k_inv_gj<R, 8, 4>), dim3(batch), dim3(NTG),
                           (gj_stage_bytes<R, 8, 4>()), s, n, s0, sw, A, sA, rb, w.used,
                           w.piv, active, status);
      } else {
        allow_lds(k_inv_gj<R, 16, 2>, gj_stage_bytes<R, 16, 2>());
        hipLaunchKernelGGL((k_inv_gj<R, 16, 2>), dim3(batch), dim3(NTG),
                           (gj_stage_bytes<R, 16, 2>()), s, n, s0, sw, A, sA, rb, w.used,
                           w.piv, active, status);
      }
    }
    HN_CHECK(hipGetLastError());
    if (rb) {  // deferred rank-sw update of every row of A outside the super-panel
      GemmArgs<R> g;
      g.m = n; g.n = n; g.k = sw;
      g.A = rb; g.sA = (long long)kInvSuper * n; g.lda = n; g.opA = OP_N;
      g.B = A + s0; g.sB = sA; g.ldb = n; g.opB = OP_N;
      g.C = A; g.sC = sA; g.ldc = n;
      g.alpha = {1.0, 0.0}; g.beta = {1.0, 0.0};
      g.active = active; g.skip_r0 = s0; g.skip_r1 = s0 + sw;
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

int inv_lds_max_n(int real_bytes) { return real_bytes == 8 ? kLuMaxN : lds_max_n<float>(); }

size_t inv_work_bytes(int n, int batch, int real_bytes) {
  size_t c = 2 * (size_t)real_bytes;
  auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
  return al((size_t)n * kInvSuper * c * batch) + al((size_t)n * batch * sizeof(int)) * 2;
}

template <typename R>
InvWork<R> inv_work_carve(void *buf, int n, int batch) {
  typedef typename InvWork<R>::CT C;
  auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
  InvWork<R> w;
  char *p = (char *)buf;
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
