// hn_kernels.h -- elementwise / reduction kernels of the engine (host side).
#pragma once
#include "hn_common.h"

namespace hn {

constexpr int kMaxCont = 10;  // MAXNCONT, lnParams.h
constexpr int kMaxBlockOrder = 1024;  // largest block the batched inverse handles
constexpr size_t kNone = ~(size_t)0;

// Where a block lives: item b reads pool + slot[b] * slot_stride + off.
struct BlkRef {
  const double2 *pool = nullptr;
  long long slot_stride = 0;
  long long off = -1;  // -1: absent
};

// out_b = z_b * S - H  (S: explicit block, identity (sid=1) or zero (sid=0, S absent))
//         - Sig_b (optional)          rows x cols, column-major, ld = rows
// z_b = E[b] + i*eta (conjz: use conj(z_b)).  trans: write the conjugate
// transpose of (z S - H) instead (used for Co and the coupling blocks).
struct FormArgs {
  BlkRef H, S;
  int s_identity = 0;
  const int *slot = nullptr;
  const double *E = nullptr;
  double eta = 0.0;
  const double *Ei = nullptr;  // optional per-item imaginary part added to eta
  int conjz = 0;
  int dagger = 0;
  int rows = 0, cols = 0;
  double2 *out = nullptr; long long s_out = 0;
  double2 *out2 = nullptr; long long s_out2 = 0;  // optional copy
  const double2 *sig = nullptr; long long s_sig = 0;
  const double2 *sig2 = nullptr; long long s_sig2 = 0;
  const double2 *sig3 = nullptr; long long s_sig3 = 0;  // e.g. an inelastic Sigma_r block
  const double2 *sig_diag = nullptr; long long s_sig_diag = 0;  // optional diagonal -= vector
  const int *active = nullptr;
  int batch = 0;
  int item_div = 1;  // item index for E/slot = b % item_div (contact items c*B + b)
};
void form_block(const FormArgs &a, hipStream_t s);

// strided batched copy dst_b = src_b (n complex per item)
void copy_batched(double2 *dst, long long sd, const double2 *src, long long ss, long long n,
                  int batch, const int *active, hipStream_t s);
void copy_batched_c(float2 *dst, long long sd, const float2 *src, long long ss, long long n,
                    int batch, const int *active, hipStream_t s);

// Decimation convergence state machine, contselfenergy.F90:300-310:
// m = max|C1_b|; if m <= acc { if ok_b: exit (act=0, ncyc=it) else ok_b=1 } else ok_b=0.
// act_b stays 1 while iterating; upd_b = 1 if the update part of this
// iteration must run.  At it == maxit items still active get status NOT_CONVERGED.
// crit 0: max |C1| (CPU path, contselfenergy.F90:300); crit 1: sum |Re|+|Im| of C1
// (GPU path, gpuroutines.cu:866-878).
void decim_check(const double2 *C1, long long sC1, int n2, int batch, double acc, int it,
                 int maxit, int *act, int *upd, int *okco, int *ncyc, int *status, int crit,
                 hipStream_t s);
void decim_check_c(const float2 *C1, long long sC1, int n2, int batch, double acc, int it,
                   int maxit, int *act, int *upd, int *okco, int *ncyc, int *status, int crit,
                   hipStream_t s);

// out[i] = (a[i] != 0) | (b[i] != 0)
void or_flags(const int *a, const int *b, int *out, int n, hipStream_t s);

// Mirror test of two contacts in every slot: flags[slot] = 1 iff
// X0 == Y0 and X1 == Y1^H bitwise (X0/Y0: PL blocks, X1/Y1: PL1->PL2 couplings),
// n x n blocks at the given offsets.  flags must be preset to 1.
void mirror_check(const double2 *pool, long long slot_stride, int nslots, int n, long long oX0,
                  long long oY0, long long oX1, long long oY1, int *flags, hipStream_t s);

// count of nonzero flags (written to *count, device)
void count_nonzero(const int *flags, int batch, int *count, hipStream_t s);

// Gam_b = i (Sig_b - Sig_b^H)
void gamma_batched(const double2 *Sig, long long sS, double2 *Gam, long long sG, int n,
                   int batch, hipStream_t s);

// T_b = | Re( Tr[Gam A] - Tr[W2 G^H] ) |,  A = i (G - G^H)
// (calculate_single_transmission_2_contacts, iterative_cpu.F90:399-470)
// written to T[b * tstride]; rmask (device, n bytes, may be null): only rows r with
// rmask[r] != 0 enter the trace (get_tun_mask, iterative_cpu.F90:585-617)
void trace_T2(const double2 *Gam, const double2 *G, const double2 *W2, long long s, int n,
              int batch, double *T, hipStream_t st, int tstride = 1,
              const unsigned char *rmask = nullptr);
// T_b = | Re Tr[W] | for W = Gam2 G21 Gam1 G21^H computed as Re sum (X)_ij conj(G21)_ij,
// written to T[b * tstride]; rmask as for trace_T2 (rows of X)
void trace_prod_conj(const double2 *X, const double2 *G, long long sX, long long sG, int rows,
                     int cols, int batch, double *T, hipStream_t st, int tstride = 1,
                     const unsigned char *rmask = nullptr);
// local DOS of a diagonal block: d_b[i] = -Im(M_b(i,i)) / pi  (iterative.F90:1907-1918)
void dos_diag(const double2 *M, long long sM, int n, int batch, double *d, long long sd,
              hipStream_t st);
// One term of diag(G^r S) (ldos_int, integrations.F90:241-246):
// out_b[r] += scale * Im(sum_c X_b(r, c) S(c, r)) for r < rows, X_b rows x cols
// (ld rows).  S(c, r) by s_kind: 0 absent (nothing added), 1 identity
// (X_b(r, r)), 2 a cols x rows block (element c + r * cols), 3 the conjugate
// of a rows x cols block (conj of element r + c * rows: S_CD = S_DC^H).
struct DiagProdArgs {
  const double2 *X = nullptr; long long sX = 0;
  int rows = 0, cols = 0;
  BlkRef S;
  int s_kind = 0;
  const int *slot = nullptr;
  double *out = nullptr; long long sout = 0;
  double scale = 0.0;
  int batch = 0;
};
void diag_prod(const DiagProdArgs &a, hipStream_t st);
// LDOS projections (iterative.F90:1926-1944): out_b[l] = sum of d_b[idx[k] - 1] over
// k in [ptr[l], ptr[l+1]) with idx[k] <= ndev (1-based, device arrays); full != 0:
// out_b[i] = d_b[i] for every device orbital
void ldos_reduce(const double *d, long long sd, int ndev, const int *ptr, const int *idx,
                 int nldos, int full, double *out, long long so, int batch, hipStream_t st);
// flag_b = 1 where max_ij |X_b(ij)| < eps (flags are only set, never cleared)
void small_flag(const double2 *X, long long sX, long long nn, double eps, int batch, int *flag,
                hipStream_t st);
// T[b * tstride] = v where flag_b != 0
void flag_set(double *T, int tstride, const int *flag, double v, int batch, hipStream_t st);
// out_b = Re Tr[X_b Y_b] (square n): sum_ij X_ij Y_ji, scaled by alpha, accumulated
void trace_prod(const double2 *X, const double2 *Y, long long sX, long long sY, int n,
                int batch, double alpha, double *out, hipStream_t st);

// extract / scatter diagonals: d_b[i] = M_b(i,i) * scale[i]  (scale may be null)
void diag_extract(const double2 *M, long long sM, int n, int batch, double2 *d, long long sd,
                  const double *scale, hipStream_t st);

// max |X_b - Y_b| over entries, per item, into err[b]
void maxabs_diff(const double2 *X, const double2 *Y, long long sX, long long sY, long long n,
                 int batch, double *err, hipStream_t st);

// element-wise fill: out_b = value (n complex)
void fill_batched(double2 *out, long long so, long long n, int batch, double2 value,
                  hipStream_t st);
void fill_int(int *out, int n, int value, hipStream_t st);
// dst item (didx ? didx[b] : b) = src item (sidx ? sidx[b] : b), n complex each;
// items whose index is negative are skipped
void copy_indexed(double2 *dst, long long sd, const int *didx, const double2 *src, long long ss,
                  const int *sidx, long long n, int batch, hipStream_t st);
// out_b = diag(v) (n x n) for every item
void fill_diag(double2 *out, long long so, int n, double2 v, int batch, hipStream_t st);
// dst_b(r0 + i, c0 + j) = src_b(i, j) for i < rows, j < cols (src column-major
// with leading dimension rows, dst with leading dimension ldd)
void place_block(double2 *dst, long long sd, int ldd, int r0, int c0, const double2 *src,
                 long long ss, int rows, int cols, int batch, hipStream_t st);
// dst[i] = src[i] (kernel copies, usable inside captured graphs)
void copy_real(double *dst, const double *src, int n, hipStream_t st);
void copy_int(int *dst, const int *src, int n, hipStream_t st);
void fill_double(double *out, int n, double value, hipStream_t st);

// precision conversion  z -> c and back, with optional active mask
void z2c(const double2 *src, long long ss, float2 *dst, long long sd, long long n, int batch,
         const int *active, hipStream_t st);
void c2z(const float2 *src, long long ss, double2 *dst, long long sd, long long n, int batch,
         const int *active, hipStream_t st);

// ---- SCBA helpers -------------------------------------------------------------
// Y_b = X_b^H  (X rows x cols -> Y cols x rows)
void dagger_batched(const double2 *X, long long sX, double2 *Y, long long sY, int rows, int cols,
                    int batch, const int *active, hipStream_t st);
// d_b[i] = M_b(i,i) * scale[i] for active items
void diag_extract_masked(const double2 *M, long long sM, int n, int batch, double2 *d,
                         long long sd, const double *scale, const int *active, hipStream_t st);
// acc[e] += sum_b w[b] X_b[e] for e < n, items summed in order b = 0..batch-1
void accumulate(double2 *acc, const double2 *X, long long sX, long long n, const double2 *w,
                int batch, hipStream_t st);
// X_b[e] = 0 where M_b[e] == 0 (n entries per item)
void mask_zero(double2 *X, long long sX, const double2 *M, long long sM, long long n, int batch,
               hipStream_t st);
// X_b *= c_b, c_b = f_ref - f_k when k != ref-1 and |f_k - f_ref| > eps, else 0
// (frm_b = frm + b * frm_stride; k 0-based, ref 1-based)
void frm_scale(double2 *X, long long sX, long long n, const double *frm, int frm_stride, int k,
               int ref, double eps, int batch, hipStream_t st);

// Sn_b = diag(v_b) + sum_q w_q(b) G_q,b   (q < nq <= 2; w_q = frm_b[j_q] - frm_b[ref-1],
// dropped when |w_q| <= eps); frm_b = frm + b * frm_stride
struct SnArgs {
  double2 *Sn = nullptr; long long sSn = 0;
  int n = 0;
  const double2 *v = nullptr; long long sv = 0;  // diagonal (may be null)
  const double2 *G[2] = {nullptr, nullptr}; long long sG = 0;
  int j[2] = {0, 0};
  int nq = 0;
  const double *frm = nullptr; int frm_stride = 0; int ref = 0;
  double eps = 0.0;
  const double2 *X = nullptr; long long sX = 0;  // optional full block added (inelastic Sigma_n)
  const int *active = nullptr;
  int batch = 0;
};
void sn_form(const SnArgs &a, hipStream_t st);

// Inelastic self-energy of one block pair for every (E,k) item (self_energy.F90:41-321
// on one process): S(iK*ne+iE)_rc (+)= sum_iQ K[|izr_r - izc_c| + ndz (iK + nk iQ)]
//   (fm G(iQ*ne + iE-shift)_rc + fp G(iQ*ne + iE+shift)_rc), G outside the grid = 0.
struct InelSigArgs {
  const double2 *G = nullptr; long long sG = 0;
  double2 *S = nullptr; long long sS = 0;
  int rows = 0, cols = 0;
  const int *izr = nullptr, *izc = nullptr;
  const double *K = nullptr; int ndz = 0, nk = 0, ne = 0;
  int shift = 0;
  double2 fm = {0.0, 0.0}, fp = {0.0, 0.0};
  // optional second term (same block, other G): its q sum is added after G's
  const double2 *G2 = nullptr;
  double2 fm2 = {0.0, 0.0}, fp2 = {0.0, 0.0};
  int accumulate = 0;
};
void inel_sigma(const InelSigArgs &a, hipStream_t st);
// err_b = max(err_b, max |X_b - Xold_b| over the mask), then Xold_b = X_b
// mask mode 0: diagonal entries; 1: entries where the slot's block `mask` != 0;
// 2: entries (r,c) where mask(r,c) != 0 or mask2(c,r) != 0 (mask2: cols x rows)
void masked_diff(const double2 *X, double2 *Xold, long long s, int rows, int cols, int mode,
                 BlkRef mask, BlkRef mask2, const int *slot, unsigned long long *err,
                 const int *active, int batch, hipStream_t st);
// per-item SCBA state after pass p: conv = p > 0 && err < tol; act = !conv && p+1 <= niter
void scba_state(const unsigned long long *err, double tol, int pass, int niter, int *act,
                int *iters, double *err_out, int batch, hipStream_t st);
// I_b,c = -Re( Tr[Gam Gn] - f_c Tr[Gam A] ), A = i(Gr - Gr^H)  (iterative.F90:382-414)
void trace_mw(const double2 *Gam, const double2 *Gn, const double2 *Gr, long long s, int n,
              const double *frm, int frm_stride, int c, int ncont, int batch, double *curr,
              hipStream_t st);
// out[i] = x[i]^2
void square(const double *x, double *out, int n, hipStream_t st);

// out_b = M_b + diag(v_b)*alpha  (add a scaled vector on the diagonal; v null: ones), n x n
void add_diag(double2 *M, long long sM, int n, const double2 *v, long long sv, double2 alpha,
              int batch, const int *active, hipStream_t st);

}  // namespace hn
