/*
 * negf_dd.c -- the oracle's two-contact transmission in double-double
 * arithmetic (about 106 significand bits), as an accuracy reference.
 *
 * TEST INFRASTRUCTURE ONLY (see negf_oracle.h).  It evaluates the same
 * algorithm as or_transmission on the fast path -- Sancho-Rubio decimation
 * with the CPU stopping rule (contselfenergy.F90:271-340), Sigma = Tlc GS Tcl
 * (:445-491, 552-566), build_ESH at real E (iterative.F90:1956-1982), the
 * gsmr sweep and G11 (iterative_cpu.F90:51-124, 172-198) and
 * T = |Re Tr[Gam A - Gam G Gam G^dag]| (iterative_cpu.F90:399-470) -- with
 * every product and sum carried in double-double, so the result is the
 * algorithm's exact value on the given double inputs up to ~1e-30 times the
 * conditioning.  Comparing a double-precision result (GPU, built-in loops,
 * MKL) with it measures that result's own rounding error.
 *
 * Double-double arithmetic: Dekker's TwoProduct (here with an FMA), Knuth's
 * TwoSum, and the accurate pairwise sum and quotient of Hida, Li and Bailey
 * ("Library for double-double and quad-double arithmetic", 2007).  Compile
 * with -ffp-contract=off: the error-free transformations must not be fused.
 *
 * Scope: square blocks of one size, contacts with no surface block and the
 * same PL size as the device blocks, two contacts, the lower one on block 1
 * (every configuration it is used on: the ribbons and the k-resolved slab).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "negf_oracle.h"

/* ---- scalar double-double ------------------------------------------------ */
typedef struct { double h, l; } dd;

static inline dd two_sum(double a, double b) {
  double s = a + b, v = s - a;
  dd r = {s, (a - (s - v)) + (b - v)};
  return r;
}
static inline dd quick_two_sum(double a, double b) {
  double s = a + b;
  dd r = {s, b - (s - a)};
  return r;
}
static inline dd dd_add(dd a, dd b) {
  dd s = two_sum(a.h, b.h), t = two_sum(a.l, b.l);
  s.l += t.h;
  s = quick_two_sum(s.h, s.l);
  s.l += t.l;
  return quick_two_sum(s.h, s.l);
}
static inline dd dd_neg(dd a) { dd r = {-a.h, -a.l}; return r; }
static inline dd dd_sub(dd a, dd b) { return dd_add(a, dd_neg(b)); }
static inline dd dd_mul(dd a, dd b) {
  double p = a.h * b.h;
  double e = fma(a.h, b.h, -p);
  e += a.h * b.l + a.l * b.h;
  return quick_two_sum(p, e);
}
static inline dd dd_mul_d(dd a, double b) {
  double p = a.h * b;
  double e = fma(a.h, b, -p);
  e += a.l * b;
  return quick_two_sum(p, e);
}
static inline dd dd_from(double x) { dd r = {x, 0.0}; return r; }
/* a / b: three-term long division */
static inline dd dd_div(dd a, dd b) {
  double q1 = a.h / b.h;
  dd r = dd_sub(a, dd_mul_d(b, q1));
  double q2 = r.h / b.h;
  r = dd_sub(r, dd_mul_d(b, q2));
  double q3 = r.h / b.h;
  dd q = quick_two_sum(q1, q2);
  return dd_add(q, dd_from(q3));
}

typedef struct { dd re, im; } zdd;

static inline zdd z_mul(zdd a, zdd b) {
  zdd r = {dd_sub(dd_mul(a.re, b.re), dd_mul(a.im, b.im)),
           dd_add(dd_mul(a.re, b.im), dd_mul(a.im, b.re))};
  return r;
}
static inline zdd z_recip(zdd a) {
  dd d = dd_add(dd_mul(a.re, a.re), dd_mul(a.im, a.im));
  dd inv = dd_div(dd_from(1.0), d);
  zdd r = {dd_mul(a.re, inv), dd_neg(dd_mul(a.im, inv))};
  return r;
}

/* ---- matrices: four planes (re hi, re lo, im hi, im lo), column-major ------ */
typedef struct { int n; double *rh, *rl, *ih, *il; } dmat;

static dmat dm_alloc(int n) {
  dmat m;
  size_t nn = (size_t)n * n;
  m.n = n;
  m.rh = (double *)calloc(4 * nn, sizeof(double));
  m.rl = m.rh + nn; m.ih = m.rl + nn; m.il = m.ih + nn;
  return m;
}
static void dm_free(dmat *m) { free(m->rh); m->rh = 0; }
static inline zdd dm_get(const dmat *m, size_t t) {
  zdd z = {{m->rh[t], m->rl[t]}, {m->ih[t], m->il[t]}};
  return z;
}
static inline void dm_set(dmat *m, size_t t, zdd z) {
  m->rh[t] = z.re.h; m->rl[t] = z.re.l; m->ih[t] = z.im.h; m->il[t] = z.im.l;
}
static void dm_copy(dmat *d, const dmat *s) {
  memcpy(d->rh, s->rh, 4 * sizeof(double) * (size_t)s->n * s->n);
}
static void dm_identity(dmat *m) {
  memset(m->rh, 0, 4 * sizeof(double) * (size_t)m->n * m->n);
  for (int i = 0; i < m->n; ++i) m->rh[i + (size_t)i * m->n] = 1.0;
}

/* z * S - H with z = zr + i zi (exact double products carried as dd);
 * S NULL: identity; conj: use z*; dag: store the conjugate transpose */
static void dm_form(dmat *m, double zr, double zi, const zc *S, const zc *H, int dag) {
  int n = m->n;
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) {
      size_t t = i + (size_t)j * n;
      double sr = S ? creal(S[t]) : (i == j ? 1.0 : 0.0);
      double si = S ? cimag(S[t]) : 0.0;
      zdd v;
      /* (zr + i zi)(sr + i si) - H */
      v.re = dd_sub(dd_sub(dd_mul(dd_from(zr), dd_from(sr)), dd_mul(dd_from(zi), dd_from(si))),
                    dd_from(H ? creal(H[t]) : 0.0));
      v.im = dd_sub(dd_add(dd_mul(dd_from(zr), dd_from(si)), dd_mul(dd_from(zi), dd_from(sr))),
                    dd_from(H ? cimag(H[t]) : 0.0));
      if (dag) {
        v.im = dd_neg(v.im);
        dm_set(m, j + (size_t)i * n, v);
      } else {
        dm_set(m, t, v);
      }
    }
}

/* y += x * b over len elements (x, y: plane columns; b a dd complex scalar).
 * The accurate dd sum of the products and of the accumulator, written out so
 * that the loop vectorises over the elements. */
static inline void zaxpy_dd(int len, const double *restrict xrh, const double *restrict xrl,
                            const double *restrict xih, const double *restrict xil, double brh,
                            double brl, double bih, double bil, double *restrict yrh,
                            double *restrict yrl, double *restrict yih, double *restrict yil) {
  for (int i = 0; i < len; ++i) {
    /* p1 = x.re b.re, p2 = x.im b.im, p3 = x.re b.im, p4 = x.im b.re */
    double p1 = xrh[i] * brh, e1 = fma(xrh[i], brh, -p1) + (xrh[i] * brl + xrl[i] * brh);
    double p2 = xih[i] * bih, e2 = fma(xih[i], bih, -p2) + (xih[i] * bil + xil[i] * bih);
    double p3 = xrh[i] * bih, e3 = fma(xrh[i], bih, -p3) + (xrh[i] * bil + xrl[i] * bih);
    double p4 = xih[i] * brh, e4 = fma(xih[i], brh, -p4) + (xih[i] * brl + xil[i] * brh);
    double s, v, e, t, w, f, s2;
    /* re: y.re + ((p1, e1) - (p2, e2)) */
    s = p1 - p2; v = s - p1; e = (p1 - (s - v)) + (-p2 - v);
    t = e1 - e2; w = t - e1; f = (e1 - (t - w)) + (-e2 - w);
    e += t; s2 = s + e; e = e - (s2 - s); e += f; s = s2 + e; e = e - (s - s2);
    {
      double ah = yrh[i], al = yrl[i];
      double S1 = ah + s, V1 = S1 - ah, E1 = (ah - (S1 - V1)) + (s - V1);
      double T1 = al + e, W1 = T1 - al, F1 = (al - (T1 - W1)) + (e - W1);
      E1 += T1; double S2 = S1 + E1; E1 = E1 - (S2 - S1); E1 += F1;
      double S3 = S2 + E1; E1 = E1 - (S3 - S2);
      yrh[i] = S3; yrl[i] = E1;
    }
    /* im: y.im + ((p3, e3) + (p4, e4)) */
    s = p3 + p4; v = s - p3; e = (p3 - (s - v)) + (p4 - v);
    t = e3 + e4; w = t - e3; f = (e3 - (t - w)) + (e4 - w);
    e += t; s2 = s + e; e = e - (s2 - s); e += f; s = s2 + e; e = e - (s - s2);
    {
      double ah = yih[i], al = yil[i];
      double S1 = ah + s, V1 = S1 - ah, E1 = (ah - (S1 - V1)) + (s - V1);
      double T1 = al + e, W1 = T1 - al, F1 = (al - (T1 - W1)) + (e - W1);
      E1 += T1; double S2 = S1 + E1; E1 = E1 - (S2 - S1); E1 += F1;
      double S3 = S2 + E1; E1 = E1 - (S3 - S2);
      yih[i] = S3; yil[i] = E1;
    }
  }
}

/* C = sa * A * B + (beta ? C : 0), sa = +-1; all n x n, C distinct from A
 * and B.  Column j of C is accumulated as a sum of columns of A scaled by
 * B(l, j), in row panels that stay in cache. */
static void dm_gemm(int sa, const dmat *A, const dmat *B, int beta, dmat *C) {
  const int n = A->n;
  const size_t nn = (size_t)n * n;
  if (!beta) memset(C->rh, 0, 4 * sizeof(double) * nn);
  const int RB = 64;
  for (int i0 = 0; i0 < n; i0 += RB) {
    const int len = i0 + RB < n ? RB : n - i0;
    for (int j = 0; j < n; ++j) {
      const size_t cj = (size_t)j * n + i0;
      for (int l = 0; l < n; ++l) {
        const size_t tb = l + (size_t)j * n;
        const double brh = sa * B->rh[tb], brl = sa * B->rl[tb];
        const double bih = sa * B->ih[tb], bil = sa * B->il[tb];
        if (brh == 0.0 && bih == 0.0 && brl == 0.0 && bil == 0.0) continue;
        const size_t al = (size_t)l * n + i0;
        zaxpy_dd(len, A->rh + al, A->rl + al, A->ih + al, A->il + al, brh, brl, bih, bil,
                 C->rh + cj, C->rl + cj, C->ih + cj, C->il + cj);
      }
    }
  }
}

/* C += s * A elementwise (s = +-1) */
static void dm_axpy(int s, const dmat *A, dmat *C) {
  size_t nn = (size_t)A->n * A->n;
  for (size_t t = 0; t < nn; ++t) {
    zdd a = dm_get(A, t), c = dm_get(C, t);
    if (s < 0) { a.re = dd_neg(a.re); a.im = dd_neg(a.im); }
    c.re = dd_add(c.re, a.re);
    c.im = dd_add(c.im, a.im);
    dm_set(C, t, c);
  }
}

/* column c of m from row r0: plane pointers */
#define COL(m, c, r0) (m)->rh + (size_t)(c) * (m)->n + (r0), (m)->rl + (size_t)(c) * (m)->n + (r0), \
                      (m)->ih + (size_t)(c) * (m)->n + (r0), (m)->il + (size_t)(c) * (m)->n + (r0)

/* Ainv = A^-1 by LU with partial pivoting (largest |Re|+|Im| of the high
 * parts) and forward/back substitution against the identity; returns 0, or
 * j + 1 for an exactly zero pivot in column j. */
static int dm_inv(const dmat *A, dmat *Ainv) {
  const int n = A->n;
  dmat LU = dm_alloc(n);
  dm_copy(&LU, A);
  int *piv = (int *)malloc(sizeof(int) * n);
  int info = 0;
  for (int j = 0; j < n; ++j) {
    int p = j;
    double best = -1.0;
    for (int i = j; i < n; ++i) {
      size_t t = i + (size_t)j * n;
      double v = fabs(LU.rh[t]) + fabs(LU.ih[t]);
      if (v > best) { best = v; p = i; }
    }
    piv[j] = p;
    if (best == 0.0) { if (!info) info = j + 1; continue; }
    if (p != j)
      for (int c = 0; c < n; ++c) {
        size_t a = j + (size_t)c * n, b = p + (size_t)c * n;
        zdd x = dm_get(&LU, a), y = dm_get(&LU, b);
        dm_set(&LU, a, y);
        dm_set(&LU, b, x);
      }
    zdd r = z_recip(dm_get(&LU, j + (size_t)j * n));
    for (int i = j + 1; i < n; ++i) {
      size_t t = i + (size_t)j * n;
      dm_set(&LU, t, z_mul(dm_get(&LU, t), r));
    }
    for (int c = j + 1; c < n; ++c) {
      zdd u = dm_get(&LU, j + (size_t)c * n);
      if (u.re.h == 0.0 && u.im.h == 0.0) continue;
      zaxpy_dd(n - j - 1, COL(&LU, j, j + 1), -u.re.h, -u.re.l, -u.im.h, -u.im.l,
               COL(&LU, c, j + 1));
    }
  }
  if (!info) {
    /* X = P I, then X <- L^-1 X, X <- U^-1 X column by column of L / U */
    dm_identity(Ainv);
    for (int j = 0; j < n; ++j)
      if (piv[j] != j)
        for (int c = 0; c < n; ++c) {
          size_t a = j + (size_t)c * n, b = piv[j] + (size_t)c * n;
          zdd x = dm_get(Ainv, a), y = dm_get(Ainv, b);
          dm_set(Ainv, a, y);
          dm_set(Ainv, b, x);
        }
    for (int c = 0; c < n; ++c)
      for (int j = 0; j < n; ++j) {
        zdd xj = dm_get(Ainv, j + (size_t)c * n);
        if (xj.re.h == 0.0 && xj.im.h == 0.0) continue;
        zaxpy_dd(n - j - 1, COL(&LU, j, j + 1), -xj.re.h, -xj.re.l, -xj.im.h, -xj.im.l,
                 COL(Ainv, c, j + 1));
      }
    zdd *rd = (zdd *)malloc(sizeof(zdd) * n);
    for (int j = 0; j < n; ++j) rd[j] = z_recip(dm_get(&LU, j + (size_t)j * n));
    for (int c = 0; c < n; ++c)
      for (int j = n - 1; j >= 0; --j) {
        size_t t = j + (size_t)c * n;
        zdd xj = z_mul(dm_get(Ainv, t), rd[j]);
        dm_set(Ainv, t, xj);
        if (xj.re.h == 0.0 && xj.im.h == 0.0) continue;
        zaxpy_dd(j, COL(&LU, j, 0), -xj.re.h, -xj.re.l, -xj.im.h, -xj.im.l, COL(Ainv, c, 0));
      }
    free(rd);
  }
  free(piv);
  dm_free(&LU);
  return info;
}

static double dm_max_abs(const dmat *m) {
  double r = 0.0;
  size_t nn = (size_t)m->n * m->n;
  for (size_t t = 0; t < nn; ++t) {
    double v = hypot(m->rh[t] + m->rl[t], m->ih[t] + m->il[t]);
    if (v > r) r = v;
  }
  return r;
}

/* Sancho-Rubio with the CPU stopping rule (max|C1| <= sgfacc on two
 * consecutive iterations, at most 300), contselfenergy.F90:271-340; Go out */
static int dd_decimation(dmat *Go, dmat *Ao, dmat *Bo, dmat *Co, double sgfacc) {
  const int n = Ao->n;
  dmat Aos = dm_alloc(n), GoXCo = dm_alloc(n), C1 = dm_alloc(n), Self = dm_alloc(n);
  dmat GoXBo = dm_alloc(n), B1 = dm_alloc(n);
  dm_copy(&Aos, Ao);
  int okCo = 0, i1;
  for (i1 = 1; i1 <= 300; ++i1) {
    dm_inv(Ao, Go);
    dm_gemm(1, Go, Co, 0, &GoXCo);
    dm_gemm(1, Co, &GoXCo, 0, &C1);
    if (dm_max_abs(&C1) <= sgfacc) {
      if (okCo) break;
      okCo = 1;
    } else {
      okCo = 0;
    }
    dm_gemm(1, Bo, &GoXCo, 0, &Self);
    dm_axpy(-1, &Self, &Aos);
    dm_axpy(-1, &Self, Ao);
    dm_gemm(1, Go, Bo, 0, &GoXBo);
    dm_gemm(1, Bo, &GoXBo, 0, &B1);
    dm_copy(Bo, &B1);
    dm_gemm(-1, Co, &GoXBo, 1, Ao);
    dm_copy(Co, &C1);
  }
  dm_inv(&Aos, Go);
  dm_free(&Aos); dm_free(&GoXCo); dm_free(&C1); dm_free(&Self); dm_free(&GoXBo); dm_free(&B1);
  return i1;
}

/* contact c of s at z = E + i delta: Sigma = Tlc GS Tcl */
static int dd_contact_sigma(const or_struct *s, int c, double E, double delta, double sgfacc,
                            dmat *Sigma) {
  const int n = s->nc[c];
  dmat Ao = dm_alloc(n), Bo = dm_alloc(n), Co = dm_alloc(n), Go = dm_alloc(n);
  dm_form(&Ao, E, delta, s->S00 ? s->S00[c] : 0, s->H00[c], 0);
  if (s->S01 && s->S01[c]) {
    dm_form(&Bo, E, delta, s->S01[c], s->H01[c], 0);
    dm_form(&Co, E, -delta, s->S01[c], s->H01[c], 1);
  } else {  /* z * 0 - H01 (a NULL S with z = 0 contributes nothing) */
    dm_form(&Bo, 0.0, 0.0, 0, s->H01[c], 0);
    dm_form(&Co, 0.0, 0.0, 0, s->H01[c], 1);
  }
  int ncyc = dd_decimation(&Go, &Ao, &Bo, &Co, sgfacc);
  dmat Tlc = dm_alloc(n), Tcl = dm_alloc(n), TG = dm_alloc(n);
  const zc *SMC = s->SMC ? s->SMC[c] : 0;
  if (SMC) {
    dm_form(&Tlc, E, delta, SMC, s->HMC[c], 0);
    dm_form(&Tcl, E, -delta, SMC, s->HMC[c], 1);
  } else {
    dm_form(&Tlc, 0.0, 0.0, 0, s->HMC[c], 0);
    dm_form(&Tcl, 0.0, 0.0, 0, s->HMC[c], 1);
  }
  dm_gemm(1, &Tlc, &Go, 0, &TG);
  dm_gemm(1, &TG, &Tcl, 0, Sigma);
  dm_free(&Ao); dm_free(&Bo); dm_free(&Co); dm_free(&Go);
  dm_free(&Tlc); dm_free(&Tcl); dm_free(&TG);
  return ncyc;
}

/* ESH blocks at real E: E S - H (S NULL -> identity / zero) */
static void dd_esh(const zc *S, const zc *H, double E, int diag, dmat *out) {
  if (S || diag) dm_form(out, E, 0.0, S, H, 0);
  else dm_form(out, 0.0, 0.0, 0, H, 0);
}

/* T(E) in double-double; T_out[0] + T_out[1] is the value, ncyc_out[ncont]
 * (may be NULL) the decimation counts.  Returns 0, or -1 for a structure
 * outside the scope above, or the zero-pivot code of an inverse. */
int dd_transmission(const or_struct *s, double E, double delta, double sgfacc, double *T_out,
                    int *ncyc_out) {
  const int nbl = s->nbl, n = s->bsz[0];
  if (s->ncont != 2 || s->ns) return -1;
  for (int i = 0; i < nbl; ++i) if (s->bsz[i] != n) return -1;
  for (int c = 0; c < 2; ++c) if (s->nc[c] != n) return -1;
  const int ct1 = s->cblk[0] < s->cblk[1] ? 0 : 1;  /* contact on the lower block */
  if (s->cblk[ct1] != 1 || nbl < 2) return -1;
  dmat Sig[2] = {dm_alloc(n), dm_alloc(n)};
  for (int c = 0; c < 2; ++c) {
    int k = dd_contact_sigma(s, c, E, delta, sgfacc, &Sig[c]);
    if (ncyc_out) ncyc_out[c] = k;
  }
  /* backward sweep: g_nbl = inv(ESH_nbl), g_i = inv(ESH_ii - ESH_i,i+1 g_i+1 ESH_i+1,i) */
  dmat D = dm_alloc(n), U = dm_alloc(n), L = dm_alloc(n), g = dm_alloc(n), w1 = dm_alloc(n);
  int info = 0;
  for (int i = nbl - 1; i >= 0; --i) {
    dd_esh(s->Sd ? s->Sd[i] : 0, s->Hd[i], E, 1, &D);
    for (int c = 0; c < 2; ++c)
      if (s->cblk[c] - 1 == i) dm_axpy(-1, &Sig[c], &D);
    if (i < nbl - 1) {
      dd_esh(s->Su ? s->Su[i] : 0, s->Hu[i], E, 0, &U);
      dd_esh(s->Sl ? s->Sl[i] : 0, s->Hl[i], E, 0, &L);
      dm_gemm(-1, &U, &g, 0, &w1);     /* -ESH_i,i+1 g_i+1 */
      dm_gemm(1, &w1, &L, 1, &D);      /* D += w1 ESH_i+1,i */
    }
    int r = dm_inv(&D, &g);
    if (r && !info) info = r;
  }
  /* g now holds G11.  T = |Re Tr[Gam A - Gam G Gam G^dag]|, A = i(G - G^dag) */
  dmat Gam = dm_alloc(n), GA = dm_alloc(n), AA = dm_alloc(n), w2 = dm_alloc(n);
  const dmat *Sg = &Sig[ct1];
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) {
      size_t t = i + (size_t)j * n, tt = j + (size_t)i * n;
      zdd a = dm_get(Sg, t), b = dm_get(Sg, tt);
      /* i (a - conj(b)) = (-(a.im + b.im), a.re - b.re) */
      zdd v = {dd_neg(dd_add(a.im, b.im)), dd_sub(a.re, b.re)};
      dm_set(&Gam, t, v);
      zdd gg = dm_get(&g, tt);
      gg.im = dd_neg(gg.im);
      dm_set(&GA, t, gg);
    }
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) {
      size_t t = i + (size_t)j * n;
      zdd a = dm_get(&g, t), b = dm_get(&GA, t);
      /* i (a - b) */
      zdd v = {dd_neg(dd_sub(a.im, b.im)), dd_sub(a.re, b.re)};
      dm_set(&AA, t, v);
    }
  dm_gemm(1, &Gam, &g, 0, &w1);
  dm_gemm(1, &w1, &Gam, 0, &w2);
  dm_gemm(1, &w2, &GA, 0, &w1);     /* Gam G Gam G^dag */
  dm_gemm(1, &Gam, &AA, 0, &w2);    /* Gam A */
  dd tr = dd_from(0.0);
  for (int i = 0; i < n; ++i) {
    size_t t = i + (size_t)i * n;
    tr = dd_add(tr, dd_sub((dd){w2.rh[t], w2.rl[t]}, (dd){w1.rh[t], w1.rl[t]}));
  }
  if (tr.h < 0) tr = dd_neg(tr);
  T_out[0] = tr.h;
  T_out[1] = tr.l;
  dm_free(&Sig[0]); dm_free(&Sig[1]);
  dm_free(&D); dm_free(&U); dm_free(&L); dm_free(&g); dm_free(&w1);
  dm_free(&Gam); dm_free(&GA); dm_free(&AA); dm_free(&w2);
  return info;
}

/* the same over a list of energies, OpenMP over points */
int dd_transmission_batch(const or_struct *s, int npts, const double *E, double delta,
                          double sgfacc, double *T_out, int *ncyc_out, int nthreads) {
  int bad = 0;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads > 0 ? nthreads : 1) reduction(| : bad)
  for (int p = 0; p < npts; ++p) {
    int r = dd_transmission(s, E[p], delta, sgfacc, T_out + 2 * p,
                            ncyc_out ? ncyc_out + s->ncont * p : 0);
    if (r) bad |= 1;
  }
  return bad ? -1 : 0;
}
