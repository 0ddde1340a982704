/*
 * hipnegf.h -- C ABI of the MI355X-native per-(E,k) Green's-function engine.
 *
 * Three layers, mirroring the three boundaries of the reference:
 *
 *  (1) Block operations: one-for-one replacements of the cu_* symbols that
 *      src/cudautils.F90:236-431 binds from Fortran (implemented in the
 *      reference by src/gpuroutines.cu).  Same argument order and meaning,
 *      column-major interleaved complex, sizes as size_t, scalars by pointer,
 *      int status return (0 = success).  Device pointers unless stated.
 *      T = complex float (C variants) or complex double (Z variants), passed
 *      as void* so no vendor complex type leaks into the ABI.
 *
 *  (2) Batched engine: entry points the reference lacks -- device-resident
 *      structures with one slot per transverse k, batched transmission and
 *      elastic-dephasing SCBA solves over many (E,k) items with per-item
 *      status words, and an explicit precision mode.
 *
 *  (3) The public libNEGF C API subset of src/api/libnegf_api.f90 for the
 *      Landauer / Meir-Wingreen / dephasing path (negf_* symbols), with the
 *      same int handler[4] convention.  Declared in hipnegf_api.h.
 *
 * All functions are thread-compatible; one engine handle must not be used
 * from two threads at once.
 */
#ifndef HIPNEGF_H
#define HIPNEGF_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Opaque context.  Stands in for both cublasHandle and cusolverDnHandle of
 * the reference (lib_param.F90:90-96, 164-165): one stream + workspace. */
typedef struct hipnegf_ctx *hipnegf_handle;

/* ------------------------------------------------------------------------ */
/* (1) block operations                                                      */
/* ------------------------------------------------------------------------ */

/* memory: cu_createMat / cu_deleteMat / cu_cudaMallocAsync / cu_cudaFreeAsync /
 * cu_copyMatH2D / cu_copyMatH2D_async / cu_copyMatD2H / cu_meminfo
 * (gpuroutines.cu:248-304, 989-996; bound at cudautils.F90:236-431).
 * delete/free null *d on success. */
int hipnegf_createMat(void **d, size_t bytes);
int hipnegf_deleteMat(void **d);
int hipnegf_mallocAsync(void **d, size_t bytes);
int hipnegf_freeAsync(void **d);
int hipnegf_copyMatH2D(void *h, void *d, size_t bytes);
int hipnegf_copyMatH2D_async(void *h, void *d, size_t bytes);
int hipnegf_copyMatD2H(void *h, void *d, size_t bytes);
int hipnegf_meminfo(size_t *free_bytes, size_t *total_bytes);
/* cudaDeviceSynchronize bound directly at cudautils.F90:261-263 */
int hipnegf_deviceSynchronize(void);

/* device / handles: cu_cudaGetDeviceCount, cu_cudaGetDevice, cu_cudaSetDevice,
 * cu_cublasInit/Finalize, cu_cusolverInit/Finalize (gpuroutines.cu:311-374).
 * Init creates a context on the current device; Finalize destroys it. */
int hipnegf_getDeviceCount(int *count);
int hipnegf_getDevice(int *device);
int hipnegf_setDevice(int device);
int hipnegf_blasInit(hipnegf_handle *h);
int hipnegf_blasFinalize(hipnegf_handle h);
int hipnegf_solverInit(hipnegf_handle *h);
int hipnegf_solverFinalize(hipnegf_handle h);

/* C = alpha op(A) op(B) + beta C;  cu_{C,Z}multMat (gpuroutines.cu:395-446):
 * dagger 0: op = N/N (A m x k, B k x n); 1: op(A) = A^H (A stored k x m);
 * 2: op(B) = B^H (B stored n x k). */
int hipnegf_CmultMat(hipnegf_handle h, size_t m, size_t n, size_t k, const void *alpha,
                     const void *A, const void *B, const void *beta, void *C, int dagger);
int hipnegf_ZmultMat(hipnegf_handle h, size_t m, size_t n, size_t k, const void *alpha,
                     const void *A, const void *B, const void *beta, void *C, int dagger);

/* Ainv = inverse(A), n x n, A unchanged; cu_{C,Z}inverse (gpuroutines.cu:449-523).
 * The reference's caller pre-sets Ainv = I (cudautils.F90:650); not required here.
 * Returns 0, or 1 if the matrix is exactly singular (no process exit).
 * n <= 1024 (the reference's ZGETRF has no limit): larger n returns -1 with
 * hipnegf_last_error set. */
int hipnegf_Cinverse(hipnegf_handle hb, hipnegf_handle hs, void *A, void *Ainv, size_t n);
int hipnegf_Zinverse(hipnegf_handle hb, hipnegf_handle hs, void *A, void *Ainv, size_t n);

/* hipnegf_Zinverse without any device allocation: no counterpart in the
 * reference.  For callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * temporaries live in one library-owned device array per GPU (5 MiB, held
 * until the process ends), so n <= 512 (larger n returns -1 with
 * hipnegf_last_error set).  Calls that overlap on one GPU share that array
 * and race; serialise them.  Same kernels and results as hipnegf_Zinverse;
 * the current device must be the handle's. */
int hipnegf_Zinverse_noalloc(hipnegf_handle hb, void *A, void *Ainv, size_t n);

/* C = alpha op(A) + beta op(B); cu_{C,Z}matsum (gpuroutines.cu:567-617):
 * dagger 0: A, B m x n; 1: op(A) = A^H (A stored n x m); 2: op(B) = B^H. */
int hipnegf_Cmatsum(hipnegf_handle h, size_t m, size_t n, const void *alpha, const void *A,
                    const void *beta, const void *B, void *C, int dagger);
int hipnegf_Zmatsum(hipnegf_handle h, size_t m, size_t n, const void *alpha, const void *A,
                    const void *beta, const void *B, void *C, int dagger);

/* A = identity (n x n); cu_{C,Z}initmat (gpuroutines.cu:619-639). */
int hipnegf_Cinitmat(void *A, size_t n);
int hipnegf_Zinitmat(void *A, size_t n);

/* sum_i Re A(i,i) [over mask]; cu_{C,Z}trace (gpuroutines.cu:642-694).
 * h_mask: host array of n int (Fortran logical, 4 bytes) -- the reference
 * declares bool here, a width mismatch with the Fortran caller; int is what
 * the Fortran side actually passes. */
float hipnegf_Ctrace(hipnegf_handle h, void *A, size_t n, void *h_mask, int mask_present);
double hipnegf_Ztrace(hipnegf_handle h, void *A, size_t n, void *h_mask, int mask_present);

/* B = A (nelem complex); cu_{C,Z}matcopy (gpuroutines.cu:697-721). */
int hipnegf_Cmatcopy(hipnegf_handle h, const void *A, void *B, size_t nelem);
int hipnegf_Zmatcopy(hipnegf_handle h, const void *A, void *B, size_t nelem);

/* *sum = sum |Re| + |Im| over n elements; cu_{C,Z}asum (gpuroutines.cu:723-739). */
int hipnegf_Casum(hipnegf_handle h, void *A, float *sum, size_t n);
int hipnegf_Zasum(hipnegf_handle h, void *A, double *sum, size_t n);

/* Sancho-Rubio decimation with HOST matrices, cu_{C,Z}decimation
 * (gpuroutines.cu:742-986, called via decimation_gpu, contselfenergy.F90:347-379):
 * Go_out = surface Green's function, Ao/Bo/Co n x n (read only here),
 * convergence as the reference GPU path: sum(|Re|+|Im|)(C1) <= SGFACC on two
 * consecutive iterations (the Fortran caller passes SGFACC*n*n), <= 300
 * iterations.  tf32 is accepted and ignored (no TF32 on CDNA4).
 * Returns 0, or 2 if not converged (Go_out then holds the last iterate). */
int hipnegf_Cdecimation(hipnegf_handle hb, hipnegf_handle hs, void *h_Go_out, void *h_Ao,
                        void *h_Bo, void *h_Co, size_t n, int tf32, int *ncyc, float SGFACC);
int hipnegf_Zdecimation(hipnegf_handle hb, hipnegf_handle hs, void *h_Go_out, void *h_Ao,
                        void *h_Bo, void *h_Co, size_t n, int tf32, int *ncyc, double SGFACC);

/* ------------------------------------------------------------------------ */
/* (2) batched engine (no counterpart in the reference)                      */
/* ------------------------------------------------------------------------ */

typedef struct hipnegf_engine_s *hipnegf_engine;
typedef struct hipnegf_struct_s *hipnegf_struct;

/* block kinds for hipnegf_struct_set_block */
enum hipnegf_block_kind {
  HIPNEGF_HD = 0,  /* H(i,i)        bsz[i] x bsz[i]     (i = 0..nbl-1)     */
  HIPNEGF_SD = 1,  /* S(i,i)                                                 */
  HIPNEGF_HU = 2,  /* H(i,i+1)      bsz[i] x bsz[i+1]   (i = 0..nbl-2)     */
  HIPNEGF_SU = 3,
  HIPNEGF_HL = 4,  /* H(i+1,i)      bsz[i+1] x bsz[i]                      */
  HIPNEGF_SL = 5,
  HIPNEGF_H00 = 6, /* contact c: PL1                nc x nc                 */
  HIPNEGF_S00 = 7,
  HIPNEGF_H01 = 8, /* contact c: PL1 -> PL2 (away from the device)          */
  HIPNEGF_S01 = 9,
  HIPNEGF_HMC = 10, /* contact c: device block cblk -> surface + PL1
                       bsz x (ns + nc)                                       */
  HIPNEGF_SMC = 11,
  /* contact c with a surface block of ns orbitals before its two PLs
   * (surface_green, contselfenergy.F90:186-239): */
  HIPNEGF_HSS = 12, /* surface                       ns x ns                 */
  HIPNEGF_SSS = 13,
  HIPNEGF_HS1 = 14, /* surface -> PL1                ns x nc                 */
  HIPNEGF_SS1 = 15,
  HIPNEGF_H1S = 16, /* PL1 -> surface                nc x ns                 */
  HIPNEGF_S1S = 17
};
/* smode bits: 1 = S(i,i) given (else identity), 2 = S(i,i+1)/S(i+1,i) given
 * (else zero), 4 = contact S blocks given (else S00 = I, S01 = SMC = 0). */
enum hipnegf_smode { HIPNEGF_S_DIAG = 1, HIPNEGF_S_OFFDIAG = 2, HIPNEGF_S_CONTACT = 4 };
enum hipnegf_precision { HIPNEGF_FP64 = 0, HIPNEGF_MIXED = 1 };
/* per-item status bits (replace process exit / __trap in the reference) */
enum hipnegf_status {
  HIPNEGF_ST_OK = 0,
  HIPNEGF_ST_SINGULAR = 1,
  HIPNEGF_ST_NOT_CONVERGED = 2,
  HIPNEGF_ST_NONFINITE = 4
};

int hipnegf_engine_create(int device, hipnegf_engine *e);
int hipnegf_engine_destroy(hipnegf_engine e);
/* last error message of the calling thread ("" if none) */
const char *hipnegf_last_error(void);

/* Device structure: nbl principal layers of sizes bsz[nbl] (the PL partition
 * of ln_structure.F90:80-150), ncont contacts attached to 1-based blocks
 * cblk[ncont] with PL size nc[ncont], nslots identical slots (e.g. one per k).
 * Every PL and every contact (surface + PL) holds at most 1024 orbitals, the
 * largest block order of the batched inverse; larger ones return -1. */
int hipnegf_struct_create(int nbl, const int *bsz, int ncont, const int *cblk, const int *nc,
                          int smode, int nslots, hipnegf_struct *st);
/* Same with contact surface blocks: ns[ncont] orbitals between the device and
 * each contact's first PL (mat_C_start - mat_B_start, ln_structure.F90:119-123;
 * 0: none).  The contact's surface Green's function is then of order ns + nc:
 * GS = inv([[z Sss - Hss, z Ss1 - Hs1], [z S1s - H1s, inv(g_PL)]])
 * (contselfenergy.F90:220-239). */
int hipnegf_struct_create_surf(int nbl, const int *bsz, int ncont, const int *cblk,
                               const int *nc, const int *ns, int smode, int nslots,
                               hipnegf_struct *st);
int hipnegf_struct_destroy(hipnegf_struct st);
/* copy one block (column-major interleaved complex double) from host memory */
int hipnegf_struct_set_block(hipnegf_struct st, int slot, int kind, int index,
                             const double *host);
/* same from device memory */
int hipnegf_struct_set_block_device(hipnegf_struct st, int slot, int kind, int index,
                                    const void *dev);

/* Options of a batched solve (zero-initialise, then set what differs). */
typedef struct {
  double delta;      /* eta added to E on the contacts (negf%delta)            */
  int ni, nf;        /* 1-based contact pair of the transmission              */
  int precision;     /* hipnegf_precision                                      */
  int max_batch;     /* 0: automatic from the HBM budget                       */
  int ns_steps;      /* mixed: FP64 Newton-Schulz steps per block inverse;
                        0: default (1), -1: none (whole sweep in FP32)        */
  int polish;        /* reserved (the decimation always runs in FP64)          */
  double eps_drop;   /* EPS drop tolerance (ln_precision), default 0           */
  int flags;         /* hipnegf_opt_flags                                      */
  int lanes;         /* concurrent item groups (own stream each); 1: one;
                        0: automatic (transmission: 2 for >= 256 items of
                        blocks >= 64; SCBA solves: 2 for >= 256 items)        */
} hipnegf_opts;
/* HIPNEGF_NO_SHARED_DECIMATION: run one Sancho-Rubio recursion per contact
 * even when two contacts are mirror images (by default such a pair shares one
 * recursion; each contact keeps its own convergence test and exit iteration).
 * HIPNEGF_GRAPH: replay the RGF sweep of a batch (the launch chain after the
 * contacts, which has no host synchronisation) from a captured hipGraph,
 * cached per structure and batch shape; results are identical. */
enum hipnegf_opt_flags {
  HIPNEGF_NO_SHARED_DECIMATION = 1,
  HIPNEGF_GRAPH = 2,
  /* surface Green's function cache (surface_green with the ln_cache memory
   * cache, contselfenergy.F90:138-173, 251-259; needs
   * hipnegf_engine_set_sgf_cache): LOAD = the reference's ReadOldSGF 0 (use
   * cached g_s when every item of a batch is cached, else compute and store;
   * loaded items report ncyc 0), STORE = ReadOldSGF 2 (compute and store). */
  HIPNEGF_SGF_LOAD = 4,
  HIPNEGF_SGF_STORE = 8,
  /* two-contact transmission with blocks of at most 16 orbitals (every block
   * the same size, no surface block, no tun_proj mask, FP64, no g_s cache)
   * runs as one launch per call, the whole solve of an item in one wave;
   * this flag keeps such calls on the batched block-operation path */
  HIPNEGF_NO_FUSED = 16
};

/* Coherent transmission T(E_i, slot_i), i < nitems (iterative.F90:1669-1780
 * per item).  Host arrays: E[nitems], slot[nitems] in; T[nitems],
 * ncyc[nitems*ncont] (decimation iterations, may be NULL), status[nitems] out. */
int hipnegf_solve_transmission(hipnegf_engine e, hipnegf_struct st, int nitems,
                               const double *E, const int *slot, const hipnegf_opts *o,
                               double *T, int *ncyc, int *status);
/* Same with all arrays in device memory (no host synchronisation of inputs). */
int hipnegf_solve_transmission_device(hipnegf_engine e, hipnegf_struct st, int nitems,
                                      const double *dE, const int *dslot,
                                      const hipnegf_opts *o, double *dT, int *dncyc,
                                      int *dstatus);

/* Several contact pairs in one solve (host arrays): the contacts' self-energies
 * and the gsmr blocks are computed once per item and shared by all pairs, as in
 * the multi-terminal branch of calculate_transmissions (iterative.F90:1716-1766).
 * ni[npairs], nf[npairs] 1-based; T[nitems*npairs] item-major.  One pair with
 * two contacts and the lower one on block 1 takes the 2-contact fast path
 * (iterative.F90:1704-1715); o->ni / o->nf are ignored. */
int hipnegf_solve_transmission_pairs(hipnegf_engine e, hipnegf_struct st, int nitems,
                                     const double *E, const int *slot, const hipnegf_opts *o,
                                     int npairs, const int *ni, const int *nf, double *T,
                                     int *ncyc, int *status);

/* Local DOS requested alongside the transmission. */
typedef struct {
  double dos_delta;  /* extra imaginary part of the device energy (negf%dos_delta)  */
  int nldos;         /* number of LDOS projections (negf%ndos_proj)                 */
  int full;          /* != 0: one value per device orbital instead of projections  */
  const int *ptr;    /* [nldos+1], ptr[0] = 0: projection l is idx[ptr[l]..ptr[l+1]-1] */
  const int *idx;    /* 1-based device orbitals (dos_proj(l)%indexes)              */
} hipnegf_dos_opts;

/* Transmission and local DOS per item (host arrays):
 * calculate_transmissions_and_dos (iterative.F90:1821-1952) as
 * tunneling_and_dos calls it (integrations.F90:1745-1766): contacts at
 * E + i delta, the device at E + i dos_delta; T[nitems*npairs] item-major by
 * the 2-contact formula on Gr(bl1,bl1) (two contacts) or the N-contact column
 * sweep; ldos[nitems*nout] item-major, nout = full ? ndev : nldos, where a
 * projection sums -Im(Gr_ii)/pi over its orbitals, skipping orbitals beyond
 * the device (iterative.F90:1926-1944).  The `full` form is the reference's
 * dos_proj(1)%indexes(1) == 0 case (iterative.F90:1930-1931).  The caller
 * applies kwght.  ncyc (may be NULL) and status as for the transmission. */
int hipnegf_solve_transmission_dos(hipnegf_engine e, hipnegf_struct st, int nitems,
                                   const double *E, const int *slot, const hipnegf_opts *o,
                                   int npairs, const int *ni, const int *nf,
                                   const hipnegf_dos_opts *d, double *T, double *ldos, int *ncyc,
                                   int *status);

/* Density-matrix integrands (the per-point work of contour_int and
 * real_axis_int, integrations.F90:844-1176).  acc (host, interleaved complex,
 * in/out) holds dense blocks packed column-major at the offsets of
 * hipnegf_density_layout: Gd[nbl] (bsz_i x bsz_i), Gu[nbl-1] = (i,i+1),
 * Gl[nbl-1] = (i+1,i), Glc[ncont] (device block cblk x contact PL, nd x nc),
 * Gcl[ncont] (nc x nd); offsets[3 nbl - 2 + 2 ncont] = total elements.
 * mode 0 (contour point E + i Ei; calculate_Gr + calculate_Gr_outer,
 * iterative.F90:90-148, 1463-1573): acc += sum_i w_i G^r(z_i), contacts at z_i.
 * mode 1 (real-axis point E; calculate_Gn_neq_components + calculate_Gn_outer,
 * iterative.F90:159-279, 1576-1665): contacts at E + i o->delta, device at E,
 * frm[nitems*(ncont+1)] occupations, ref 1-based (ncont+1: none), acc +=
 * sum_i w_i G^n(E_i).  w[2*nitems] complex weights.  outer: 0 device blocks
 * only, 1 + Glc, 2 + Glc and Gcl.  Points are summed in order.  The caller
 * masks the blocks to the overlap pattern (blk2csr, iterative.F90:1375-1450).
 * status[nitems] (may be NULL) per-item status words. */
int hipnegf_density_layout(hipnegf_struct st, long long *offsets);
int hipnegf_density_accumulate(hipnegf_engine e, hipnegf_struct st, int nitems, const double *Er,
                               const double *Ei, const int *slot, const hipnegf_opts *o, int mode,
                               const double *w, const double *frm, int ref, int outer,
                               double *acc, int *status);

/* tun_proj (set_tun_indexes, libnegf.F90:1044-1058): only the listed 1-based
 * device orbitals of the traced block enter every transmission trace of this
 * structure (get_tun_mask, iterative_cpu.F90:585-617).  n = 0 clears it. */
int hipnegf_struct_set_tun_proj(hipnegf_struct st, int n, const int *idx);

/* Elastic dephasing SCBA + Meir-Wingreen current per item (host arrays):
 * calculate_elastic_scba (iterative.F90:289-363) with the ElPhonDephD model
 * (elphdd.F90; coupling per device orbital, squared internally), then
 * iterative_meir_wingreen (iterative.F90:382-414).
 * frm[ncont+1] occupations (frm[ref-1] is the reference, ref = ncont+1 for
 * none) -- or per-item occupations if frm_per_item != 0 (frm[nitems*(ncont+1)]).
 * Out: curr[nitems*ncont], iters[nitems], err[nitems], status[nitems];
 * gn_diag / gr_diag [nitems*ndev] complex (may be NULL). */
int hipnegf_solve_scba_mw(hipnegf_engine e, hipnegf_struct st, int nitems, const double *E,
                          const int *slot, const hipnegf_opts *o, const double *coupling,
                          int niter, double tol, const double *frm, int frm_per_item, int ref,
                          double *curr, int *iters, double *err, int *status, double *gn_diag,
                          double *gr_diag);

/* meir_wingreen with an inelastic electron-phonon model (host arrays):
 * the outer SCBA of integrations.F90:1804-1966 over the (E,k) points of a
 * uniform grid E[ne] (ne >= 2) and the first nk structure slots (k weights
 * kweights[nk]).  Each pass runs calculate_elastic_scba per point
 * (coupling[ndev] elastic dephasing, may be NULL; niter_ela / tol_ela) with
 * the inelastic self-energies of the previous pass (ElPhonInel add_sigma_r /
 * add_sigma_n, elphinel.F90:540-621), then rebuilds them (compute_Sigma_n /
 * compute_Sigma_r, elphinel.F90:729-1100; selfenergy, self_energy.F90:41-321):
 *   Sigma(E,k)_rc = sum_q K[|iz_r - iz_c| + ndz (k + nk q)]
 *                   (fm G(E - shift dE, q)_rc + fp G(E + shift dE, q)_rc)
 * on the diagonal blocks (and (i,i+1), (i+1,i) if tridiag), G outside the grid
 * = 0, with (fm, fp) = (Nq, Nq+1) on G^n for Sigma_n and (Nq+1, Nq) on G^r
 * plus (i/2, -i/2) on G^n for Sigma_r.  frm[ne*(ncont+1)] occupations per
 * energy (frm[iE*(ncont+1) + ref-1] the reference).  Stops when
 * max_c |J_c - J_c(previous)| < tol (check_J_convergence, scba.F90:150-173),
 * J_c = g_spin * integrate_el_meir_wingreen(curr_mat(:,c)) (integration 0
 * trapezoidal, 1 Simpson 1/3, 2 Simpson 3/8), or after pass niter.
 * K[ndz*nk*nk]: hipnegf_inel_kmat_po / _nonpo; iz[ndev]: z bin of each device
 * orbital (bin differences must stay below ndz).  Equal PL and contact sizes.
 * Out: curr_mat[ne*ncont] (iE + ne*c, sum_k kweights[k] I_c(E,k)),
 * currents[ncont], passes, err (last max |dJ|), status[ne*nk] and
 * iters[ne*nk] (item k*ne + iE; may be NULL). */
int hipnegf_solve_inelastic_mw(hipnegf_engine e, hipnegf_struct st, int ne, const double *E,
                               int nk, const double *kweights, const hipnegf_opts *o,
                               const double *coupling, int niter_ela, double tol_ela,
                               const double *frm, int ref, const double *K, int ndz, const int *iz,
                               int shift, double Nq, int tridiag, int niter, double tol,
                               int integration, double g_spin, double *curr_mat, double *currents,
                               int *passes, double *err, int *status, int *iters);
/* K tables of the polar-optical (prepare_POKmat, elphinel.F90:318-406) and
 * non-polar optical (prepare_NonPOKmat, :411-519) models, without equivalent
 * k-points: K[iz + ndz*(iK + nk*iQ)] for z = iz*dz.  kfrac[3*nk] fractional
 * k-points, latt[9] the lattice vectors column-wise (all 0: k = 0).  Host only. */
int hipnegf_inel_kmat_po(int ndz, double dz, int nk, const double *kfrac, const double *kweights,
                         const double *latt, double coupling, double wq, double cell_area,
                         double eps0, double eps_inf, double q0, double *K);
int hipnegf_inel_kmat_nonpo(int ndz, double dz, int nk, const double *kfrac,
                            const double *kweights, const double *latt, double coupling, double D0,
                            double *K);

/* Batched kernels for tests and callers with their own drivers (device
 * pointers, item b at X + b*strideX elements): */
int hipnegf_zgemm_batched(hipnegf_engine e, int m, int n, int k, const double *alpha,
                          const void *A, long long sA, int lda, int opA, const void *B,
                          long long sB, int ldb, int opB, const double *beta, void *C,
                          long long sC, int ldc, int batch);
int hipnegf_cgemm_batched(hipnegf_engine e, int m, int n, int k, const double *alpha,
                          const void *A, long long sA, int lda, int opA, const void *B,
                          long long sB, int ldb, int opB, const double *beta, void *C,
                          long long sC, int ldc, int batch);
/* Ainv_b = inverse(A_b); A is overwritten (working matrix). status[batch] out (device). */
int hipnegf_zinverse_batched(hipnegf_engine e, int n, void *A, long long sA, void *Ainv,
                             long long sAinv, int batch, int *dstatus);
int hipnegf_cinverse_batched(hipnegf_engine e, int n, void *A, long long sA, void *Ainv,
                             long long sAinv, int batch, int *dstatus);
int hipnegf_engine_sync(hipnegf_engine e);
/* stream the engine launches on (a hipStream_t) */
void *hipnegf_engine_stream(hipnegf_engine e);

/* ---- synthetic structures (no counterpart in the reference) ------------- */
/* Fill `slot` of a structure (uniform PL size n, contact 1 on the last block,
 * contact 2 on block 1; smode 7 with overlap, else 0) with the k-resolved
 * random slab of `seed` at transverse k = (kx, ky), generated on the device. */
int hipnegf_struct_generate_kslab(hipnegf_engine e, hipnegf_struct st, int slot,
                                  unsigned long long seed, double kx, double ky, int with_overlap);
/* copy one block of a slot back to the host (same kinds as set_block) */
int hipnegf_struct_get_block(hipnegf_struct st, int slot, int kind, int index, double *host);

/* ---- serving loop (no counterpart in the reference) ----------------------
 * A persistent request loop: `workers` threads, each with its own engine and
 * HIP stream, serve transmission requests.  A request is admitted when its
 * device footprint fits the HBM budget (FCFS), its energies are cut into
 * chunks of bounded work, and the workers interleave chunks of all admitted
 * requests (policy 0: oldest request first; 1: least GPU time so far), so
 * requests that arrive while others run start between chunks.  A request's
 * results do not depend on the mix it ran in. */
typedef struct hipnegf_server_s *hipnegf_server;
typedef struct {
  int workers;                /* 1..16; 0: 2 */
  int policy;                 /* 0 FCFS, 1 least attained service */
  int max_chunk;              /* items per chunk cap; 0: automatic */
  int reserved;
  double chunk_flops;         /* target algorithmic flops per chunk; 0: 2e12 */
  long long mem_budget;       /* bytes; 0: 85 % of free HBM */
  long long worker_workspace; /* bytes per worker; 0: automatic */
  long long sgf_cache;        /* bytes of the surface-function cache the workers share
                                 (requests then load cached g_s, else compute and store
                                 them: ReadOldSGF = 0); 0: none */
  int keep_structures;        /* 1: generated structures stay resident after their
                                 requests (least recently used evicted when admission
                                 needs the memory); 0: freed when their last request ends */
  int reserved2;
} hipnegf_server_opts;
int hipnegf_server_create(int device, const hipnegf_server_opts *o, hipnegf_server *s);
/* finishes the admitted requests, drops queued ones, joins the workers */
int hipnegf_server_destroy(hipnegf_server s);
/* request on a generated slab (built in HBM at its first admission; kept
 * resident for later requests with the same arguments when keep_structures) */
int hipnegf_server_submit_kslab(hipnegf_server s, unsigned long long seed, int n, int nbl,
                                double kx, double ky, int with_overlap, int nE, const double *E,
                                const hipnegf_opts *o, unsigned long long *id);
/* request on a caller-owned resident structure (must outlive the request) */
int hipnegf_server_submit(hipnegf_server s, hipnegf_struct st, int slot, int nE, const double *E,
                          const hipnegf_opts *o, unsigned long long *id);
int hipnegf_server_poll(hipnegf_server s, unsigned long long id, int *done);
/* blocks until done; T[nE], ncyc[nE*ncont], status[nE] (any may be NULL);
 * times[4] = {queue s, latency s, GPU s, chunks}.  Returns nonzero with the
 * request's error message if it failed.  The id is released. */
int hipnegf_server_wait(hipnegf_server s, unsigned long long id, double *T, int *ncyc, int *status,
                        double *times);
/* {submitted, completed, running, queued} */
int hipnegf_server_stats(hipnegf_server s, long long *out);
/* {requests served on an already resident generated structure, resident
 * generated structures, surface functions loaded from the cache, stored} */
int hipnegf_server_cache_stats(hipnegf_server s, long long *out);

/* Device memory for the engine's surface Green's function cache (0: off).
 * Entries are keyed by (structure and its contents version, slot, E, eta,
 * contact); least recently used entries are replaced. */
int hipnegf_engine_set_sgf_cache(hipnegf_engine e, long long bytes);
/* cache hits (items x contacts loaded) and stores since the engine started */
int hipnegf_engine_sgf_stats(hipnegf_engine e, long long *hits, long long *stores);
/* Fictitious contact (pnegf%cont(c)%FictCont, contselfenergy.F90:175-185):
 * g_s = -i pi dos I instead of the decimation; contact is 0-based. */
int hipnegf_struct_set_fictitious(hipnegf_struct st, int contact, int fictitious, double dos);

/* FP64 GEMM tile choice (no counterpart in the reference): -1 the autotuner
 * (default: the first launch of each shape class times the 64 x 64, the
 * 64 x 128 and the 32 x 32 single-wave tiles and keeps the fixed rule's tile
 * unless another is 3 % faster; HIPNEGF_GEMM_TUNE=0 in the environment
 * switches it off), 0 always 64 x 64, 1 always 64 x 128, 2 always 32 x 32.
 * The tilings give bitwise the same results. */
int hipnegf_gemm_force_variant(int variant);
/* {shape classes tuned, of which chose the 64 x 128 tile, tuner on, of which
 * chose the 32 x 32 tile} */
int hipnegf_gemm_tune_stats(long long *out);

/* Launch timing (no counterpart in the reference; measurement hook for
 * benchmarks).  While enabled, every tagged launch records a HIP event pair on
 * its stream.  Tags: 0 masked GEMMs (decimation, SCBA), 1 RGF-sweep GEMMs,
 * 2 blocked-inverse trailing updates, 3 inverse panels, 4 small-block inverses,
 * 5 fused small-block transmission solves.
 * hipnegf_profile_read waits for the events of `tag` and returns the summed
 * duration (ms), algorithmic flops (-1 if any launch is not counted) and
 * algorithmic bytes, and the launch count. */
int hipnegf_profile_enable(int on);
int hipnegf_profile_read(int tag, double *ms, double *flops, double *bytes, long long *launches);
/* the same over the launches of `tag` served by one kernel variant: for the
 * GEMM tags 0 the 64 x 64 FP64 tile (k_gemm<double, tag>), 1 the 64 x 128 FP64
 * tile (k_gemm_w<tag>), 2 the FP32 tile (k_gemm<float, tag>), 3 the 32 x 32
 * single-wave FP64 tile (k_gemm_s<tag>); 0 for the others */
int hipnegf_profile_read_variant(int tag, int variant, double *ms, double *flops, double *bytes,
                                 long long *launches);
int hipnegf_profile_clear(void);
/* FP64 matrix-core issue loop with register operands only (no counterpart in
 * the reference; calibration of hardware counters): `blocks` workgroups of
 * four waves, each running eight independent chains of `iters` instructions
 * of form 0 (v_mfma_f64_4x4x4) or 1 (v_mfma_f64_16x16x4); *tflops = the
 * achieved rate of the second of two launches. */
int hipnegf_calib_mfma(int form, int iters, int blocks, double *tflops);

#ifdef __cplusplus
}
#endif

#endif /* HIPNEGF_H */
