// hn_kernels.hip -- elementwise and per-item reduction kernels of the engine.
//
// These replace the reference's per-block helper calls (zsumdns, zdagger,
// zspectral, ztrace, cu_Zmatsum / cu_Ztrace / cu_Zasum, sparsekit_drv.F90 and
// gpuroutines.cu:567-739) by whole-batch launches; all are HBM-streaming.
#include <algorithm>

#include "hn_kernels.h"

namespace hn {

namespace {

constexpr int NT = 256;

__device__ inline double2 cm(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

template <typename T> __device__ inline T wave_sum(T v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off);
  return v;
}
template <typename T> __device__ inline T wave_max(T v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    T o = __shfl_xor(v, off);
    v = (o > v || o != o) ? o : v;  // propagate NaN
  }
  return v;
}

// block-wide reductions for NT threads
template <typename T> __device__ inline T block_sum(T v, T *sh) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) sh[w] = v;
  __syncthreads();
  T r = 0;
  if (threadIdx.x == 0)
    for (int i = 0; i < NT / 64; ++i) r += sh[i];
  __syncthreads();
  return r;  // valid in thread 0
}
template <typename T> __device__ inline T block_max(T v, T *sh) {
  v = wave_max(v);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) sh[w] = v;
  __syncthreads();
  T r = 0;
  if (threadIdx.x == 0) {
    r = sh[0];
    for (int i = 1; i < NT / 64; ++i) r = (sh[i] > r || sh[i] != sh[i]) ? sh[i] : r;
  }
  __syncthreads();
  return r;
}

__global__ __launch_bounds__(NT) void k_form(FormArgs a) {
  const int b = blockIdx.y;
  if (a.active && a.active[b] == 0) return;
  const long long n = (long long)a.rows * a.cols;
  const long long e = (long long)blockIdx.x * NT + threadIdx.x;
  if (e >= n) return;
  const int r = (int)(e % a.rows), c = (int)(e / a.rows);
  const int sl = a.slot ? a.slot[b] : 0;
  double2 z = make_double2(a.E ? a.E[b] : 0.0, a.Ei ? a.eta + a.Ei[b] : a.eta);
  if (a.conjz) z.y = -z.y;
  // source element (rs, cs) of the (possibly transposed) source block
  const int rs = a.dagger ? c : r, cs = a.dagger ? r : c;
  const int srows = a.dagger ? a.cols : a.rows;
  const long long se = rs + (long long)cs * srows;
  const double2 *Hb = a.H.pool + (long long)sl * a.H.slot_stride + a.H.off;
  double2 h = Hb[se];
  double2 sv = make_double2(0.0, 0.0);
  if (a.S.off >= 0) {
    const double2 *Sb = a.S.pool + (long long)sl * a.S.slot_stride + a.S.off;
    sv = Sb[se];
  } else if (a.s_identity && rs == cs) {
    sv = make_double2(1.0, 0.0);
  }
  double2 v = cm(z, sv);
  v.x -= h.x;
  v.y -= h.y;
  if (a.dagger) v.y = -v.y;
  if (a.sig) {
    double2 sg = a.sig[(long long)b * a.s_sig + e];
    v.x -= sg.x;
    v.y -= sg.y;
  }
  if (a.sig2) {
    double2 sg = a.sig2[(long long)b * a.s_sig2 + e];
    v.x -= sg.x;
    v.y -= sg.y;
  }
  if (a.sig3) {
    double2 sg = a.sig3[(long long)b * a.s_sig3 + e];
    v.x -= sg.x;
    v.y -= sg.y;
  }
  if (a.sig_diag && r == c) {
    double2 sg = a.sig_diag[(long long)b * a.s_sig_diag + r];
    v.x -= sg.x;
    v.y -= sg.y;
  }
  a.out[(long long)b * a.s_out + e] = v;
  if (a.out2) a.out2[(long long)b * a.s_out2 + e] = v;
}

template <typename C>
__global__ __launch_bounds__(NT) void k_copy(C *dst, long long sd, const C *src, long long ss,
                                             long long n, const int *active) {
  const int b = blockIdx.y;
  if (active && active[b] == 0) return;
  const long long e = (long long)blockIdx.x * NT + threadIdx.x;
  if (e >= n) return;
  dst[(long long)b * sd + e] = src[(long long)b * ss + e];
}

template <typename C>
__global__ __launch_bounds__(NT) void k_decim_check(const C *C1, long long sC1, int n2,
                                                    double acc, int it, int maxit, int *act,
                                                    int *upd, int *okco, int *ncyc,
                                                    int *status, int crit) {
  __shared__ double sh[NT / 64];
  const int b = blockIdx.x;
  if (act[b] == 0) {
    if (threadIdx.x == 0) upd[b] = 0;
    return;
  }
  const C *X = C1 + (long long)b * sC1;
  double m = 0.0;
  if (crit == 0) {
    for (int e = threadIdx.x; e < n2; e += NT) {
      C v = X[e];
      double a = hypot((double)v.x, (double)v.y);
      m = (a > m || a != a) ? a : m;
    }
    m = block_max(m, sh);
  } else {
    for (int e = threadIdx.x; e < n2; e += NT) {
      C v = X[e];
      m += fabs((double)v.x) + fabs((double)v.y);
    }
    m = block_sum(m, sh);
  }
  if (threadIdx.x == 0) {
    if (m != m) {  // non-finite iterate: stop this item, flag it
      act[b] = 0;
      upd[b] = 0;
      ncyc[b] = it;
      status[b] |= ST_NONFINITE;
      return;
    }
    if (m <= acc) {
      if (okco[b]) {
        act[b] = 0;
        upd[b] = 0;
        ncyc[b] = it;
        return;
      }
      okco[b] = 1;
    } else {
      okco[b] = 0;
    }
    upd[b] = 1;
    if (it >= maxit) {  // the DO loop ends after this update: ncyc = maxit + 1
      ncyc[b] = maxit + 1;
      status[b] |= ST_NOT_CONVERGED;
    }
  }
}

__global__ __launch_bounds__(NT) void k_or_flags(const int *a, const int *b, int *out, int n) {
  const int i = blockIdx.x * NT + threadIdx.x;
  if (i < n) out[i] = (a[i] != 0) | (b[i] != 0);
}

__global__ __launch_bounds__(NT) void k_mirror(const double2 *pool, long long slot_stride, int n,
                                               long long oX0, long long oY0, long long oX1,
                                               long long oY1, int *flags) {
  const int slot = blockIdx.y;
  const long long e = (long long)blockIdx.x * NT + threadIdx.x;
  if (e >= (long long)n * n) return;
  const int r = (int)(e % n), c = (int)(e / n);
  const double2 *base = pool + (long long)slot * slot_stride;
  const double2 x0 = base[oX0 + e], y0 = base[oY0 + e];
  const double2 x1 = base[oX1 + e], y1t = base[oY1 + c + (long long)r * n];
  const bool ok = x0.x == y0.x && x0.y == y0.y && x1.x == y1t.x && x1.y == -y1t.y;
  if (!ok) atomicAnd(&flags[slot], 0);
}

__global__ __launch_bounds__(NT) void k_count(const int *flags, int batch, int *count) {
  __shared__ int sh[NT / 64];
  int c = 0;
  for (int i = threadIdx.x; i < batch; i += NT) c += flags[i] != 0;
  c = block_sum(c, sh);
  if (threadIdx.x == 0) *count = c;
}

__global__ __launch_bounds__(NT) void k_gamma(const double2 *Sig, long long sS, double2 *Gam,
                                              long long sG, int n) {
  const int b = blockIdx.y;
  const long long e = (long long)blockIdx.x * NT + threadIdx.x;
  if (e >= (long long)n * n) return;
  const int r = (int)(e % n), c = (int)(e / n);
  const double2 *S = Sig + (long long)b * sS;
  double2 a = S[e], t = S[c + (long long)r * n];
  // i (a - conj(t)) = i((a.x - t.x) + i(a.y + t.y)) = -(a.y + t.y) + i (a.x - t.x)
  Gam[(long long)b * sG + e] = make_double2(-(a.y + t.y), a.x - t.x);
}

__global__ __launch_bounds__(NT) void k_trace_T2(const double2 *Gam, const double2 *G,
                                                 const double2 *W2, long long s, int n,
                                                 double *T, int tstride,
                                                 const unsigned char *rmask) {
  __shared__ double sh[NT / 64];
  const int b = blockIdx.x;
  const double2 *Gm = Gam + (long long)b * s, *Gb = G + (long long)b * s, *Wb = W2 + (long long)b * s;
  double acc = 0.0;
  const long long nn = (long long)n * n;
  for (long long e = threadIdx.x; e < nn; e += NT) {
    const int r = (int)(e % n), c = (int)(e / n);
    if (rmask && !rmask[r]) continue;  // entry (r,c) feeds diagonal entry r of the product
    const double2 g_rc = Gb[e], g_cr = Gb[c + (long long)r * n];
    // A_cr = i (G_cr - conj(G_rc))
    const double2 d = make_double2(g_cr.x - g_rc.x, g_cr.y + g_rc.y);
    const double2 A_cr = make_double2(-d.y, d.x);
    const double2 gm = Gm[e];
    acc += gm.x * A_cr.x - gm.y * A_cr.y;          // Re(Gam_rc A_cr)
    const double2 w = Wb[e];
    acc -= w.x * g_rc.x + w.y * g_rc.y;            // Re(W_rc conj(G_rc))
  }
  acc = block_sum(acc, sh);
  if (threadIdx.x == 0) T[(long long)b * tstride] = fabs(acc);
}

__global__ __launch_bounds__(NT) void k_trace_prod_conj(const double2 *X, const double2 *G,
                                                        long long sX, long long sG, int rows,
                                                        long long nn, double *T, int tstride,
                                                        const unsigned char *rmask) {
  __shared__ double sh[NT / 64];
  const int b = blockIdx.x;
  const double2 *Xb = X + (long long)b * sX, *Gb = G + (long long)b * sG;
  double acc = 0.0;
  for (long long e = threadIdx.x; e < nn; e += NT) {
    if (rmask && !rmask[e % rows]) continue;
    const double2 x = Xb[e], g = Gb[e];
    acc += x.x * g.x + x.y * g.y;
  }
  acc = block_sum(acc, sh);
  if (threadIdx.x == 0) T[(long long)b * tstride] = fabs(acc);
}

// flag_b |= (max |X_b| < eps)
__global__ __launch_bounds__(NT) void k_small_flag(const double2 *X, long long sX, long long nn,
                                                   double eps, int *flag) {
  __shared__ double sh[NT / 64];
  const int b = blockIdx.x;
  double m = 0.0;
  for (long long e = threadIdx.x; e < nn; e += NT) {
    const double2 x = X[(long long)b * sX + e];
    const double a = hypot(x.x, x.y);
    m = (a > m || a != a) ? a : m;
  }
  m = block_max(m, sh);
  if (threadIdx.x == 0 && m < eps) flag[b] = 1;
}

__global__ __launch_bounds__(NT) void k_flag_set(double *T, int tstride, const int *flag,
                                                 double v, int batch) {
  const int b = blockIdx.x * NT + threadIdx.x;
  if (b < batch && flag[b]) T[(long long)b * tstride] = v;
}

__global__ __launch_bounds__(NT) void k_trace_prod(const double2 *X, const double2 *Y,
                                                   long long sX, long long sY, int n,
                                                   double alpha, double *out) {
  __shared__ double sh[NT / 64];
  const int b = blockIdx.x;
  const double2 *Xb = X + (long long)b * sX, *Yb = Y + (long long)b * sY;
  double acc = 0.0;
  const long long nn = (long long)n * n;
  for (long long e = threadIdx.x; e < nn; e += NT) {
    const int r = (int)(e % n), c = (int)(e / n);
    const double2 x = Xb[e], y = Yb[c + (long long)r * n];
    acc += x.x * y.x - x.y * y.y;
  }
  acc = block_sum(acc, sh);
  if (threadIdx.x == 0) out[b] += alpha * acc;
}

__global__ __launch_bounds__(NT) void k_diag_extract(const double2 *M, long long sM, int n,
                                                     double2 *d, long long sd,
                                                     const double *scale) {
  const int b = blockIdx.y;
  const int i = blockIdx.x * NT + threadIdx.x;
  if (i >= n) return;
  double2 v = M[(long long)b * sM + i + (long long)i * n];
  double s = scale ? scale[i] : 1.0;
  d[(long long)b * sd + i] = make_double2(v.x * s, v.y * s);
}

__global__ __launch_bounds__(NT) void k_maxabs_diff(const double2 *X, const double2 *Y,
                                                    long long sX, long long sY, long long n,
                                                    double *err) {
  __shared__ double sh[NT / 64];
  const int b = blockIdx.x;
  double m = 0.0;
  for (long long e = threadIdx.x; e < n; e += NT) {
    double2 x = X[(long long)b * sX + e], y = Y[(long long)b * sY + e];
    double a = hypot(x.x - y.x, x.y - y.y);
    m = (a > m || a != a) ? a : m;
  }
  m = block_max(m, sh);
  if (threadIdx.x == 0) err[b] = m;
}

__global__ __launch_bounds__(NT) void k_fill(double2 *out, long long so, long long n,
                                             double2 v) {
  const int b = blockIdx.y;
  const long long e = (long long)blockIdx.x * NT + threadIdx.x;
  if (e < n) out[(long long)b * so + e] = v;
}
__global__ __launch_bounds__(NT) void k_fill_int(int *out, int n, int v) {
  const int e = blockIdx.x * NT + threadIdx.x;
  if (e < n) out[e] = v;
}
__global__ __launch_bounds__(NT) void k_fill_double(double *out, int n, double v) {
  const int e = blockIdx.x * NT + threadIdx.x;
  if (e < n) out[e] = v;
}

__global__ __launch_bounds__(NT) void k_z2c(const double2 *src, long long ss, float2 *dst,
                                            long long sd, long long n, const int *active) {
  const int b = blockIdx.y;
  if (active && active[b] == 0) return;
  const long long e = (long long)blockIdx.x * NT + threadIdx.x;
  if (e >= n) return;
  double2 v = src[(long long)b * ss + e];
  dst[(long long)b * sd + e] = make_float2((float)v.x, (float)v.y);
}
__global__ __launch_bounds__(NT) void k_c2z(const float2 *src, long long ss, double2 *dst,
                                            long long sd, long long n, const int *active) {
  const int b = blockIdx.y;
  if (active && active[b] == 0) return;
  const long long e = (long long)blockIdx.x * NT + threadIdx.x;
  if (e >= n) return;
  float2 v = src[(long long)b * ss + e];
  dst[(long long)b * sd + e] = make_double2((double)v.x, (double)v.y);
}

__global__ __launch_bounds__(NT) void k_add_diag(double2 *M, long long sM, int n,
                                                 const double2 *v, long long sv, double2 alpha,
                                                 const int *active) {
  const int b = blockIdx.y;
  if (active && active[b] == 0) return;
  const int i = blockIdx.x * NT + threadIdx.x;
  if (i >= n) return;
  const double2 x = v ? v[(long long)b * sv + i] : make_double2(1.0, 0.0);  // v == null: ones
  double2 *p = M + (long long)b * sM + i + (long long)i * n;
  double2 m = *p;
  m.x += alpha.x * x.x - alpha.y * x.y;
  m.y += alpha.x * x.y + alpha.y * x.x;
  *p = m;
}


__global__ __launch_bounds__(NT) void k_dagger(const double2 *X, long long sX, double2 *Y,
                                               long long sY, int rows, int cols,
                                               const int *active) {
  const int b = blockIdx.y;
  if (active && active[b] == 0) return;
  const long long e = (long long)blockIdx.x * NT + threadIdx.x;
  if (e >= (long long)rows * cols) return;
  const int r = (int)(e % cols), c = (int)(e / cols);  // Y is cols x rows
  const double2 x = X[(long long)b * sX + c + (long long)r * rows];
  Y[(long long)b * sY + e] = make_double2(x.x, -x.y);
}

__global__ __launch_bounds__(NT) void k_diag_extract_m(const double2 *M, long long sM, int n,
                                                       double2 *d, long long sd,
                                                       const double *scale, const int *active) {
  const int b = blockIdx.y;
  if (active && active[b] == 0) return;
  const int i = blockIdx.x * NT + threadIdx.x;
  if (i >= n) return;
  double2 v = M[(long long)b * sM + i + (long long)i * n];
  double s = scale ? scale[i] : 1.0;
  d[(long long)b * sd + i] = make_double2(v.x * s, v.y * s);
}

__global__ __launch_bounds__(NT) void k_sn_form(SnArgs a) {
  const int b = blockIdx.y;
  if (a.active && a.active[b] == 0) return;
  const long long e = (long long)blockIdx.x * NT + threadIdx.x;
  if (e >= (long long)a.n * a.n) return;
  const int r = (int)(e % a.n), c = (int)(e / a.n);
  double2 out = make_double2(0.0, 0.0);
  if (a.v && r == c) out = a.v[(long long)b * a.sv + r];
  const double *f = a.frm + (long long)b * a.frm_stride;
  for (int q = 0; q < a.nq; ++q) {
    const double w = f[a.j[q]] - f[a.ref - 1];
    if (fabs(w) > a.eps) {
      const double2 g = a.G[q][(long long)b * a.sG + e];
      out.x += w * g.x;
      out.y += w * g.y;
    }
  }
  if (a.X) {
    const double2 x = a.X[(long long)b * a.sX + e];
    out.x += x.x;
    out.y += x.y;
  }
  a.Sn[(long long)b * a.sSn + e] = out;
}

// one thread per (element, energy); the k and q sums run inside.  NKT > 0:
// the nk <= NKT shifted-energy sums of each q are formed once and held in
// registers for every k; NKT == 0: any nk, re-read per k (from L2).
// Per output: acc = S or 0; acc += K(k,q) sb1(q) for q in order; then
// acc += K(k,q) sb2(q) for q in order (the order of two separate passes).
template <int NKT>
__global__ __launch_bounds__(NT) void k_inel_sigma(InelSigArgs a) {
  const int iE = blockIdx.y;
  const long long e = (long long)blockIdx.x * NT + threadIdx.x;
  if (e >= (long long)a.rows * a.cols) return;
  const int r = (int)(e % a.rows), c = (int)(e / a.rows);
  const int dz = abs(a.izr[r] - a.izc[c]);
  const bool hm = iE - a.shift >= 0, hp = iE + a.shift < a.ne;
  auto shifted = [&](const double2 *G, double2 fm, double2 fp, int q) {
    double2 sb = make_double2(0.0, 0.0);
    if (hm) sb = cm(fm, G[(long long)(q * a.ne + iE - a.shift) * a.sG + e]);
    if (hp) {
      const double2 t = cm(fp, G[(long long)(q * a.ne + iE + a.shift) * a.sG + e]);
      sb.x += t.x;
      sb.y += t.y;
    }
    return sb;
  };
  const double *Kz = a.K + dz;
  if constexpr (NKT > 0) {
    double2 s1[NKT], s2[NKT];
#pragma unroll
    for (int q = 0; q < NKT; ++q)
      if (q < a.nk) {
        s1[q] = shifted(a.G, a.fm, a.fp, q);
        s2[q] = a.G2 ? shifted(a.G2, a.fm2, a.fp2, q) : make_double2(0.0, 0.0);
      }
    for (int k = 0; k < a.nk; ++k) {
      double2 *Sp = a.S + (long long)(k * a.ne + iE) * a.sS + e;
      double2 acc = a.accumulate ? *Sp : make_double2(0.0, 0.0);
#pragma unroll
      for (int q = 0; q < NKT; ++q)
        if (q < a.nk) {
          const double kv = Kz[(long long)a.ndz * (k + (long long)a.nk * q)];
          acc.x += kv * s1[q].x;
          acc.y += kv * s1[q].y;
        }
      if (a.G2) {
#pragma unroll
        for (int q = 0; q < NKT; ++q)
          if (q < a.nk) {
            const double kv = Kz[(long long)a.ndz * (k + (long long)a.nk * q)];
            acc.x += kv * s2[q].x;
            acc.y += kv * s2[q].y;
          }
      }
      *Sp = acc;
    }
  } else {
    for (int k = 0; k < a.nk; ++k) {
      double2 *Sp = a.S + (long long)(k * a.ne + iE) * a.sS + e;
      double2 acc = a.accumulate ? *Sp : make_double2(0.0, 0.0);
      for (int t = 0; t < (a.G2 ? 2 : 1); ++t)
        for (int q = 0; q < a.nk; ++q) {
          const double2 sb = t == 0 ? shifted(a.G, a.fm, a.fp, q) : shifted(a.G2, a.fm2, a.fp2, q);
          const double kv = Kz[(long long)a.ndz * (k + (long long)a.nk * q)];
          acc.x += kv * sb.x;
          acc.y += kv * sb.y;
        }
      *Sp = acc;
    }
  }
}

__global__ __launch_bounds__(NT) void k_masked_diff(const double2 *X, double2 *Xold, long long s,
                                                    int rows, int cols, int mode, BlkRef mask,
                                                    BlkRef mask2, const int *slot,
                                                    unsigned long long *err, const int *active) {
  __shared__ double sh[NT / 64];
  const int b = blockIdx.y;
  if (active && active[b] == 0) return;
  const long long nn = (long long)rows * cols;
  const double2 *Xb = X + (long long)b * s;
  double2 *Ob = Xold + (long long)b * s;
  const long long sl = slot ? slot[b] : 0;
  const double2 *Mb = mode ? mask.pool + sl * mask.slot_stride + mask.off : nullptr;
  const double2 *M2 = mode == 2 ? mask2.pool + sl * mask2.slot_stride + mask2.off : nullptr;
  double m = 0.0;
  for (long long e = (long long)blockIdx.x * NT + threadIdx.x; e < nn; e += (long long)gridDim.x * NT) {
    const int r = (int)(e % rows), c = (int)(e / rows);
    bool in = false;
    if (mode == 0) in = (r == c);
    else if (mode == 1) { const double2 v = Mb[e]; in = (v.x != 0.0 || v.y != 0.0); }
    else {
      const double2 v = Mb[e], w = M2[c + (long long)r * cols];
      in = (v.x != 0.0 || v.y != 0.0 || w.x != 0.0 || w.y != 0.0);
    }
    const double2 x = Xb[e];
    if (in) {
      const double2 o = Ob[e];
      const double d = hypot(x.x - o.x, x.y - o.y);
      m = (d > m || d != d) ? d : m;
    }
    Ob[e] = x;
  }
  m = block_max(m, sh);
  if (threadIdx.x == 0) {
    // non-negative doubles order like their bit patterns; NaN maps above +inf
    unsigned long long bits = (m != m) ? 0x7ff8000000000000ULL : (unsigned long long)__double_as_longlong(m);
    atomicMax(&err[b], bits);
  }
}

__global__ __launch_bounds__(NT) void k_scba_state(const unsigned long long *err, double tol,
                                                   int pass, int niter, int *act, int *iters,
                                                   double *err_out, int batch) {
  const int b = blockIdx.x * NT + threadIdx.x;
  if (b >= batch || act[b] == 0) return;
  const double e = __longlong_as_double((long long)err[b]);
  bool conv = false;
  if (pass > 0) {
    err_out[b] = e;
    conv = e < tol;
  }
  iters[b] = pass + 1;
  act[b] = (!conv && pass + 1 <= niter) ? 1 : 0;
}

__global__ __launch_bounds__(NT) void k_trace_mw(const double2 *Gam, const double2 *Gn,
                                                 const double2 *Gr, long long s, int n,
                                                 const double *frm, int frm_stride, int c,
                                                 int ncont, double *curr) {
  __shared__ double sh[NT / 64];
  const int b = blockIdx.x;
  const double2 *Gm = Gam + (long long)b * s, *Nb = Gn + (long long)b * s, *Rb = Gr + (long long)b * s;
  const double f = frm[(long long)b * frm_stride + c];
  double tg = 0.0, ta = 0.0;
  const long long nn = (long long)n * n;
  for (long long e = threadIdx.x; e < nn; e += NT) {
    const int r = (int)(e % n), cc = (int)(e / n);
    const double2 g = Gm[e];
    const double2 gn = Nb[cc + (long long)r * n];          // Gn(c, r)
    tg += g.x * gn.x - g.y * gn.y;
    const double2 a1 = Rb[cc + (long long)r * n], a2 = Rb[e];  // Gr(c,r), Gr(r,c)
    // A(c,r) = i (Gr(c,r) - conj(Gr(r,c)))
    const double2 d = make_double2(a1.x - a2.x, a1.y + a2.y);
    const double2 A = make_double2(-d.y, d.x);
    ta += g.x * A.x - g.y * A.y;
  }
  const double v = block_sum(tg - f * ta, sh);
  if (threadIdx.x == 0) curr[(long long)b * ncont + c] = -v;
}

__global__ __launch_bounds__(NT) void k_copy_indexed(double2 *dst, long long sd, const int *didx,
                                                     const double2 *src, long long ss,
                                                     const int *sidx, long long n) {
  const int b = blockIdx.y;
  const long long di = didx ? didx[b] : b, si = sidx ? sidx[b] : b;
  if (di < 0 || si < 0) return;
  for (long long e = (long long)blockIdx.x * NT + threadIdx.x; e < n; e += (long long)gridDim.x * NT)
    dst[di * sd + e] = src[si * ss + e];
}

__global__ __launch_bounds__(NT) void k_fill_diag(double2 *out, long long so, int n, double2 v) {
  const int b = blockIdx.y;
  const long long e = (long long)blockIdx.x * NT + threadIdx.x;
  if (e >= (long long)n * n) return;
  const int r = (int)(e % n), c = (int)(e / n);
  out[(long long)b * so + e] = (r == c) ? v : make_double2(0.0, 0.0);
}

__global__ __launch_bounds__(NT) void k_place(double2 *dst, long long sd, int ldd, int r0, int c0,
                                              const double2 *src, long long ss, int rows,
                                              int cols) {
  const int b = blockIdx.y;
  const long long e = (long long)blockIdx.x * NT + threadIdx.x;
  if (e >= (long long)rows * cols) return;
  const int r = (int)(e % rows), c = (int)(e / rows);
  dst[(long long)b * sd + (r0 + r) + (long long)(c0 + c) * ldd] = src[(long long)b * ss + e];
}

template <typename T>
__global__ __launch_bounds__(NT) void k_copy1(T *dst, const T *src, int n) {
  const int i = blockIdx.x * NT + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

__global__ __launch_bounds__(NT) void k_square(const double *x, double *out, int n) {
  const int i = blockIdx.x * NT + threadIdx.x;
  if (i < n) out[i] = x[i] * x[i];
}

inline dim3 grid2(long long n, int batch) {
  return dim3((unsigned)((n + NT - 1) / NT), (unsigned)batch);
}

}  // namespace

void form_block(const FormArgs &a, hipStream_t s) {
  if (a.batch <= 0 || a.rows <= 0 || a.cols <= 0) return;
  hipLaunchKernelGGL(k_form, grid2((long long)a.rows * a.cols, a.batch), dim3(NT), 0, s, a);
  HN_CHECK(hipGetLastError());
}

void copy_batched(double2 *dst, long long sd, const double2 *src, long long ss, long long n,
                  int batch, const int *active, hipStream_t s) {
  if (batch <= 0 || n <= 0) return;
  hipLaunchKernelGGL(k_copy<double2>, grid2(n, batch), dim3(NT), 0, s, dst, sd, src, ss, n, active);
  HN_CHECK(hipGetLastError());
}
void copy_batched_c(float2 *dst, long long sd, const float2 *src, long long ss, long long n,
                    int batch, const int *active, hipStream_t s) {
  if (batch <= 0 || n <= 0) return;
  hipLaunchKernelGGL(k_copy<float2>, grid2(n, batch), dim3(NT), 0, s, dst, sd, src, ss, n, active);
  HN_CHECK(hipGetLastError());
}

void decim_check(const double2 *C1, long long sC1, int n2, int batch, double acc, int it,
                 int maxit, int *act, int *upd, int *okco, int *ncyc, int *status, int crit,
                 hipStream_t s) {
  hipLaunchKernelGGL(k_decim_check<double2>, dim3(batch), dim3(NT), 0, s, C1, sC1, n2, acc, it,
                     maxit, act, upd, okco, ncyc, status, crit);
  HN_CHECK(hipGetLastError());
}
void decim_check_c(const float2 *C1, long long sC1, int n2, int batch, double acc, int it,
                   int maxit, int *act, int *upd, int *okco, int *ncyc, int *status, int crit,
                   hipStream_t s) {
  hipLaunchKernelGGL(k_decim_check<float2>, dim3(batch), dim3(NT), 0, s, C1, sC1, n2, acc, it,
                     maxit, act, upd, okco, ncyc, status, crit);
  HN_CHECK(hipGetLastError());
}

void or_flags(const int *a, const int *b, int *out, int n, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_or_flags, dim3((n + NT - 1) / NT), dim3(NT), 0, s, a, b, out, n);
  HN_CHECK(hipGetLastError());
}

void mirror_check(const double2 *pool, long long slot_stride, int nslots, int n, long long oX0,
                  long long oY0, long long oX1, long long oY1, int *flags, hipStream_t s) {
  hipLaunchKernelGGL(k_mirror, grid2((long long)n * n, nslots), dim3(NT), 0, s, pool, slot_stride,
                     n, oX0, oY0, oX1, oY1, flags);
  HN_CHECK(hipGetLastError());
}

void count_nonzero(const int *flags, int batch, int *count, hipStream_t s) {
  hipLaunchKernelGGL(k_count, dim3(1), dim3(NT), 0, s, flags, batch, count);
  HN_CHECK(hipGetLastError());
}

void gamma_batched(const double2 *Sig, long long sS, double2 *Gam, long long sG, int n,
                   int batch, hipStream_t s) {
  hipLaunchKernelGGL(k_gamma, grid2((long long)n * n, batch), dim3(NT), 0, s, Sig, sS, Gam, sG, n);
  HN_CHECK(hipGetLastError());
}

void trace_T2(const double2 *Gam, const double2 *G, const double2 *W2, long long s, int n,
              int batch, double *T, hipStream_t st, int tstride, const unsigned char *rmask) {
  hipLaunchKernelGGL(k_trace_T2, dim3(batch), dim3(NT), 0, st, Gam, G, W2, s, n, T, tstride,
                     rmask);
  HN_CHECK(hipGetLastError());
}

void trace_prod_conj(const double2 *X, const double2 *G, long long sX, long long sG, int rows,
                     int cols, int batch, double *T, hipStream_t st, int tstride,
                     const unsigned char *rmask) {
  hipLaunchKernelGGL(k_trace_prod_conj, dim3(batch), dim3(NT), 0, st, X, G, sX, sG, rows,
                     (long long)rows * cols, T, tstride, rmask);
  HN_CHECK(hipGetLastError());
}

// d_b[i] = Im((-1/pi) M_b(i,i)): the value zrconcatm_csr ('I', s = -1/pi) stores
__global__ __launch_bounds__(NT) void k_dos_diag(const double2 *M, long long sM, int n, double *d,
                                                 long long sd) {
  const int b = blockIdx.y;
  const int i = blockIdx.x * NT + threadIdx.x;
  if (i >= n) return;
  const double2 v = M[(long long)b * sM + i + (long long)i * n];
  const double c = -1.0 / 3.14159265358979323846;
  d[(long long)b * sd + i] = c * v.y + 0.0 * v.x;
}

void dos_diag(const double2 *M, long long sM, int n, int batch, double *d, long long sd,
              hipStream_t st) {
  hipLaunchKernelGGL(k_dos_diag, dim3((n + NT - 1) / NT, batch), dim3(NT), 0, st, M, sM, n, d, sd);
  HN_CHECK(hipGetLastError());
}

__global__ __launch_bounds__(NT) void k_diag_prod(DiagProdArgs a) {
  const int b = blockIdx.y;
  const int r = blockIdx.x * NT + threadIdx.x;
  if (r >= a.rows) return;
  const double2 *X = a.X + (long long)b * a.sX;
  double im = 0.0;
  if (a.s_kind == 1) {
    im = X[r + (long long)r * a.rows].y;
  } else {
    const int sl = a.slot ? a.slot[b] : 0;
    const double2 *Sb = a.S.pool + (long long)sl * a.S.slot_stride + a.S.off;
    for (int c = 0; c < a.cols; ++c) {
      const double2 x = X[r + (long long)c * a.rows];
      double2 v;
      if (a.s_kind == 2) {
        v = Sb[c + (long long)r * a.cols];
      } else {
        v = Sb[r + (long long)c * a.rows];
        v.y = -v.y;
      }
      im += x.x * v.y + x.y * v.x;
    }
  }
  a.out[(long long)b * a.sout + r] += a.scale * im;
}

void diag_prod(const DiagProdArgs &a, hipStream_t st) {
  if (a.s_kind == 0 || a.rows <= 0 || a.batch <= 0) return;
  if (a.s_kind != 1 && a.S.off < 0) throw std::runtime_error("diag_prod: S block absent");
  hipLaunchKernelGGL(k_diag_prod, dim3((a.rows + NT - 1) / NT, a.batch), dim3(NT), 0, st, a);
  HN_CHECK(hipGetLastError());
}

__global__ __launch_bounds__(NT) void k_accumulate(double2 *acc, const double2 *X, long long sX,
                                                   long long n, const double2 *w, int batch) {
  const long long e = (long long)blockIdx.x * NT + threadIdx.x;
  if (e >= n) return;
  double2 a = acc[e];
  for (int b = 0; b < batch; ++b) {
    const double2 x = X[(long long)b * sX + e], c = w[b];
    const double2 p = make_double2(c.x * x.x - c.y * x.y, c.x * x.y + c.y * x.x);
    a.x += p.x;
    a.y += p.y;
  }
  acc[e] = a;
}

void accumulate(double2 *acc, const double2 *X, long long sX, long long n, const double2 *w,
                int batch, hipStream_t st) {
  if (n <= 0 || batch <= 0) return;
  hipLaunchKernelGGL(k_accumulate, dim3((unsigned)((n + NT - 1) / NT)), dim3(NT), 0, st, acc, X, sX, n,
                     w, batch);
  HN_CHECK(hipGetLastError());
}

__global__ __launch_bounds__(NT) void k_mask_zero(double2 *X, long long sX, const double2 *M,
                                                  long long sM, long long n) {
  const int b = blockIdx.y;
  const long long e = (long long)blockIdx.x * NT + threadIdx.x;
  if (e >= n) return;
  const double2 m = M[(long long)b * sM + e];
  if (m.x == 0.0 && m.y == 0.0) X[(long long)b * sX + e] = make_double2(0.0, 0.0);
}

void mask_zero(double2 *X, long long sX, const double2 *M, long long sM, long long n, int batch,
               hipStream_t st) {
  if (n <= 0 || batch <= 0) return;
  hipLaunchKernelGGL(k_mask_zero, grid2(n, batch), dim3(NT), 0, st, X, sX, M, sM, n);
  HN_CHECK(hipGetLastError());
}

__global__ __launch_bounds__(NT) void k_frm_scale(double2 *X, long long sX, long long n,
                                                  const double *frm, int frm_stride, int k, int ref,
                                                  double eps) {
  const int b = blockIdx.y;
  const long long e = (long long)blockIdx.x * NT + threadIdx.x;
  if (e >= n) return;
  const double *f = frm + (long long)b * frm_stride;
  const double fr = f[ref - 1], fk = f[k];
  const double c = (k != ref - 1 && fabs(fk - fr) > eps) ? fr - fk : 0.0;
  double2 &x = X[(long long)b * sX + e];
  x = make_double2(c * x.x, c * x.y);
}

void frm_scale(double2 *X, long long sX, long long n, const double *frm, int frm_stride, int k,
               int ref, double eps, int batch, hipStream_t st) {
  if (n <= 0 || batch <= 0) return;
  hipLaunchKernelGGL(k_frm_scale, grid2(n, batch), dim3(NT), 0, st, X, sX, n, frm, frm_stride, k,
                     ref, eps);
  HN_CHECK(hipGetLastError());
}

// one thread per (item, projection): the list summed in its own order from 0
__global__ __launch_bounds__(NT) void k_ldos_reduce(const double *d, long long sd, int ndev,
                                                    const int *ptr, const int *idx, int nldos,
                                                    int full, double *out, long long so,
                                                    int batch) {
  const long long t = (long long)blockIdx.x * NT + threadIdx.x;
  const int nout = full ? ndev : nldos;
  if (t >= (long long)nout * batch) return;
  const int b = (int)(t / nout), l = (int)(t % nout);
  const double *db = d + (long long)b * sd;
  double acc = 0.0;
  if (full) {
    acc += db[l];
  } else {
    for (int k = ptr[l]; k < ptr[l + 1]; ++k) {
      const int i2 = idx[k];
      if (i2 >= 1 && i2 <= ndev) acc += db[i2 - 1];  // callers reject i2 < 1
    }
  }
  out[(long long)b * so + l] = acc;
}

void ldos_reduce(const double *d, long long sd, int ndev, const int *ptr, const int *idx,
                 int nldos, int full, double *out, long long so, int batch, hipStream_t st) {
  const long long tot = (long long)(full ? ndev : nldos) * batch;
  if (tot <= 0) return;
  hipLaunchKernelGGL(k_ldos_reduce, dim3((unsigned)((tot + NT - 1) / NT)), dim3(NT), 0, st, d, sd,
                     ndev, ptr, idx, nldos, full, out, so, batch);
  HN_CHECK(hipGetLastError());
}
void small_flag(const double2 *X, long long sX, long long nn, double eps, int batch, int *flag,
                hipStream_t st) {
  hipLaunchKernelGGL(k_small_flag, dim3(batch), dim3(NT), 0, st, X, sX, nn, eps, flag);
  HN_CHECK(hipGetLastError());
}
void flag_set(double *T, int tstride, const int *flag, double v, int batch, hipStream_t st) {
  hipLaunchKernelGGL(k_flag_set, dim3((batch + NT - 1) / NT), dim3(NT), 0, st, T, tstride, flag,
                     v, batch);
  HN_CHECK(hipGetLastError());
}

void trace_prod(const double2 *X, const double2 *Y, long long sX, long long sY, int n,
                int batch, double alpha, double *out, hipStream_t st) {
  hipLaunchKernelGGL(k_trace_prod, dim3(batch), dim3(NT), 0, st, X, Y, sX, sY, n, alpha, out);
  HN_CHECK(hipGetLastError());
}

void diag_extract(const double2 *M, long long sM, int n, int batch, double2 *d, long long sd,
                  const double *scale, hipStream_t st) {
  hipLaunchKernelGGL(k_diag_extract, grid2(n, batch), dim3(NT), 0, st, M, sM, n, d, sd, scale);
  HN_CHECK(hipGetLastError());
}

void maxabs_diff(const double2 *X, const double2 *Y, long long sX, long long sY, long long n,
                 int batch, double *err, hipStream_t st) {
  hipLaunchKernelGGL(k_maxabs_diff, dim3(batch), dim3(NT), 0, st, X, Y, sX, sY, n, err);
  HN_CHECK(hipGetLastError());
}

void fill_batched(double2 *out, long long so, long long n, int batch, double2 value,
                  hipStream_t st) {
  if (n <= 0 || batch <= 0) return;
  hipLaunchKernelGGL(k_fill, grid2(n, batch), dim3(NT), 0, st, out, so, n, value);
  HN_CHECK(hipGetLastError());
}
void fill_int(int *out, int n, int value, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_fill_int, dim3((n + NT - 1) / NT), dim3(NT), 0, st, out, n, value);
  HN_CHECK(hipGetLastError());
}

void copy_indexed(double2 *dst, long long sd, const int *didx, const double2 *src, long long ss,
                  const int *sidx, long long n, int batch, hipStream_t st) {
  if (batch <= 0 || n <= 0) return;
  const unsigned gx = (unsigned)std::min<long long>((n + NT - 1) / NT, 64);
  hipLaunchKernelGGL(k_copy_indexed, dim3(gx, batch), dim3(NT), 0, st, dst, sd, didx, src, ss, sidx, n);
  HN_CHECK(hipGetLastError());
}
void fill_diag(double2 *out, long long so, int n, double2 v, int batch, hipStream_t st) {
  if (batch <= 0) return;
  hipLaunchKernelGGL(k_fill_diag, grid2((long long)n * n, batch), dim3(NT), 0, st, out, so, n, v);
  HN_CHECK(hipGetLastError());
}
void place_block(double2 *dst, long long sd, int ldd, int r0, int c0, const double2 *src,
                 long long ss, int rows, int cols, int batch, hipStream_t st) {
  if (batch <= 0 || rows <= 0 || cols <= 0) return;
  hipLaunchKernelGGL(k_place, grid2((long long)rows * cols, batch), dim3(NT), 0, st, dst, sd, ldd,
                     r0, c0, src, ss, rows, cols);
  HN_CHECK(hipGetLastError());
}
void copy_real(double *dst, const double *src, int n, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_copy1<double>, dim3((n + NT - 1) / NT), dim3(NT), 0, st, dst, src, n);
  HN_CHECK(hipGetLastError());
}
void copy_int(int *dst, const int *src, int n, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_copy1<int>, dim3((n + NT - 1) / NT), dim3(NT), 0, st, dst, src, n);
  HN_CHECK(hipGetLastError());
}
void fill_double(double *out, int n, double value, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_fill_double, dim3((n + NT - 1) / NT), dim3(NT), 0, st, out, n, value);
  HN_CHECK(hipGetLastError());
}

void z2c(const double2 *src, long long ss, float2 *dst, long long sd, long long n, int batch,
         const int *active, hipStream_t st) {
  if (n <= 0 || batch <= 0) return;
  hipLaunchKernelGGL(k_z2c, grid2(n, batch), dim3(NT), 0, st, src, ss, dst, sd, n, active);
  HN_CHECK(hipGetLastError());
}
void c2z(const float2 *src, long long ss, double2 *dst, long long sd, long long n, int batch,
         const int *active, hipStream_t st) {
  if (n <= 0 || batch <= 0) return;
  hipLaunchKernelGGL(k_c2z, grid2(n, batch), dim3(NT), 0, st, src, ss, dst, sd, n, active);
  HN_CHECK(hipGetLastError());
}

void dagger_batched(const double2 *X, long long sX, double2 *Y, long long sY, int rows, int cols,
                    int batch, const int *active, hipStream_t st) {
  hipLaunchKernelGGL(k_dagger, grid2((long long)rows * cols, batch), dim3(NT), 0, st, X, sX, Y, sY,
                     rows, cols, active);
  HN_CHECK(hipGetLastError());
}
void diag_extract_masked(const double2 *M, long long sM, int n, int batch, double2 *d,
                         long long sd, const double *scale, const int *active, hipStream_t st) {
  hipLaunchKernelGGL(k_diag_extract_m, grid2(n, batch), dim3(NT), 0, st, M, sM, n, d, sd, scale,
                     active);
  HN_CHECK(hipGetLastError());
}
void inel_sigma(const InelSigArgs &a, hipStream_t st) {
  if (a.ne <= 0 || a.rows <= 0 || a.cols <= 0) return;
  const dim3 g = grid2((long long)a.rows * a.cols, a.ne);
  if (a.nk <= 1) hipLaunchKernelGGL(k_inel_sigma<1>, g, dim3(NT), 0, st, a);
  else if (a.nk <= 2) hipLaunchKernelGGL(k_inel_sigma<2>, g, dim3(NT), 0, st, a);
  else if (a.nk <= 4) hipLaunchKernelGGL(k_inel_sigma<4>, g, dim3(NT), 0, st, a);
  else if (a.nk <= 8) hipLaunchKernelGGL(k_inel_sigma<8>, g, dim3(NT), 0, st, a);
  else if (a.nk <= 16) hipLaunchKernelGGL(k_inel_sigma<16>, g, dim3(NT), 0, st, a);
  else hipLaunchKernelGGL(k_inel_sigma<0>, g, dim3(NT), 0, st, a);
  HN_CHECK(hipGetLastError());
}
void sn_form(const SnArgs &a, hipStream_t st) {
  hipLaunchKernelGGL(k_sn_form, grid2((long long)a.n * a.n, a.batch), dim3(NT), 0, st, a);
  HN_CHECK(hipGetLastError());
}
void masked_diff(const double2 *X, double2 *Xold, long long s, int rows, int cols, int mode,
                 BlkRef mask, BlkRef mask2, const int *slot, unsigned long long *err,
                 const int *active, int batch, hipStream_t st) {
  long long nn = (long long)rows * cols;
  unsigned gx = (unsigned)std::min<long long>((nn + NT - 1) / NT, 64);
  hipLaunchKernelGGL(k_masked_diff, dim3(gx, batch), dim3(NT), 0, st, X, Xold, s, rows, cols, mode,
                     mask, mask2, slot, err, active);
  HN_CHECK(hipGetLastError());
}
void scba_state(const unsigned long long *err, double tol, int pass, int niter, int *act,
                int *iters, double *err_out, int batch, hipStream_t st) {
  hipLaunchKernelGGL(k_scba_state, dim3((batch + NT - 1) / NT), dim3(NT), 0, st, err, tol, pass,
                     niter, act, iters, err_out, batch);
  HN_CHECK(hipGetLastError());
}
void trace_mw(const double2 *Gam, const double2 *Gn, const double2 *Gr, long long s, int n,
              const double *frm, int frm_stride, int c, int ncont, int batch, double *curr,
              hipStream_t st) {
  hipLaunchKernelGGL(k_trace_mw, dim3(batch), dim3(NT), 0, st, Gam, Gn, Gr, s, n, frm, frm_stride,
                     c, ncont, curr);
  HN_CHECK(hipGetLastError());
}
void square(const double *x, double *out, int n, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_square, dim3((n + NT - 1) / NT), dim3(NT), 0, st, x, out, n);
  HN_CHECK(hipGetLastError());
}

void add_diag(double2 *M, long long sM, int n, const double2 *v, long long sv, double2 alpha,
              int batch, const int *active, hipStream_t st) {
  hipLaunchKernelGGL(k_add_diag, grid2(n, batch), dim3(NT), 0, st, M, sM, n, v, sv, alpha, active);
  HN_CHECK(hipGetLastError());
}

}  // namespace hn
