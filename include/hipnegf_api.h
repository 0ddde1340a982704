/*
 * hipnegf_api.h -- the libNEGF public C API subset of the per-energy path,
 * implemented on the MI355X engine (libhipnegf.so).
 *
 * Same entry-point names, argument order and meaning as the Fortran-generated
 * C API of the reference (src/api/libnegf_api.f90; handler convention
 * src/api/libnegfAPICommon.f90:34-39; parameter block src/api/lnParams.h), so
 * a C/C++/ctypes caller of the reference links against this library
 * unchanged for the Landauer / Meir-Wingreen / elastic-dephasing path.
 * Also provided: the local DOS alongside T and the density matrix (contour
 * and real-axis integration of the same block solver).  Entry points outside
 * that path (k-point meshes, inelastic models, the UPT reader) are not
 * provided.
 *
 * Errors: the reference stops the process (error stop); here an entry point
 * that fails records a message (negf_get_last_error) and leaves its outputs
 * untouched; negf_get_status returns it as an int (0 = ok).
 */
#ifndef HIPNEGF_API_H
#define HIPNEGF_API_H

#ifdef __cplusplus
#include <cstdbool>
extern "C" {
#else
#include <stdbool.h>
#endif

#define MAXNCONT 10
#define MAXNLAYERS 10000
#define NEGF_HSIZE 4 /* DAC_handlerSize, libnegfAPICommon.f90:39 */

/* identical layout to src/api/lnParams.h:15-55 */
struct lnparams {
  int verbose;
  int readold_dm_sgfs;
  int readold_t_sgfs;
  int sgfcache;
  int spin;
  int kpoint;
  double g_spin;
  double delta;
  int deltaModel;
  double wmax;
  double dos_delta;
  double eneconv;
  double kwght;
  double ec;
  double ev;
  double deltaec;
  double deltaev;
  double emin;
  double emax;
  double estep;
  double estep_coarse;
  double mu_n[MAXNCONT];
  double mu_p[MAXNCONT];
  double mu[MAXNCONT];
  double contact_dos[MAXNCONT];
  bool fictcont[MAXNCONT];
  double kbt_dm[MAXNCONT];
  double kbt_t[MAXNCONT];
  double scba_inelastic_tol;
  double scba_elastic_tol;
  int np_n[2];
  int np_p[2];
  int np_real;
  int n_kt;
  int n_poles;
  int ni[MAXNLAYERS];
  int nf[MAXNLAYERS];
  char dore[1];
  int min_or_max;
  bool is_s_is;
};

/* session (libnegf_api.f90:62-134, 633-666) */
void negf_gethandlersize(int *handlerSize);
void negf_init_session(int *handler);
void negf_init(int *handler);
void negf_destruct_libnegf(int *handler);
void negf_destruct_session(int *handler);
void negf_set_verbosity(int *handler, int verbose_lev);

/* structure (libnegf_api.f90:146-185, 272-285) -- 1-based indices */
void negf_init_contacts(int *handler, int ncont);
void negf_init_structure(int *handler, int ncont, int *surfstart, int *surfend, int *contend,
                         int npl, int *plend, int *cblk);

/* parameters (libnegf_api.f90:338-365) */
/* Structure helpers (libnegf_api.f90:190-269, 102-116): negf_contact_blocks
 * finds, from the rows surfstart..contend of H, the device PL each contact
 * couples to (find_cblocks, libnegf.F90:2172-2234; an error when two);
 * negf_get_pls returns npl and, with copy != 0, plend, ncont and cblk;
 * negf_get_version prints the library version. */
void negf_contact_blocks(int *handler, int ncont, int *surfstart, int *surfend, int *contend,
                         int npl, int *plend, int *cblk);
void negf_get_pls(int *handler, int *npl, int *ncont, int *plend, int *cblk, int copy);
void negf_get_version(int *handler);
/* Diagnostics (libnegf_api.f90:545-567, 1188-1200, 1274-1286, 1462-1475):
 * negf_print_mat writes H.dat / S.dat into the scratch path in printcsr's
 * text format (2i8, 2f20.10 per stored entry); negf_write_partition,
 * negf_print_tnegf and negf_mem_stats print the partition, the parameters
 * and the memory in use to stdout. */
void negf_print_mat(int *handler);
void negf_write_partition(int *handler);
void negf_print_tnegf(int *handler);
void negf_mem_stats(int *handler);
/* min_or_max of the density paths (libnegf_api.f90:1172-1185) */
void negf_set_reference(int *handler, int *minmax);
void negf_set_params(int *handler, struct lnparams *params);
void negf_get_params(int *handler, struct lnparams *params);

/* H and S in CSR; 0-based if rowpnt[0] == 0, else 1-based (libnegf.F90:303-434).
 * A: interleaved complex double (same bytes as double _Complex / complex(8)).
 * iK: 1-based k index (negf_create_hs sets the count). */
void negf_create_hs(int *handler, int nHk);
void negf_set_h(int *handler, int nrow, const double *A, int *JA, int *IA, int iK);
void negf_set_s(int *handler, int nrow, const double *A, int *JA, int *IA, int iK);
void negf_set_h_cp(int *handler, int nrow, double *a_re, double *a_im, int *ja, int *ia);
void negf_set_s_cp(int *handler, int nrow, double *a_re, double *a_im, int *ja, int *ia);
void negf_set_s_id(int *handler, int nrow, int iK);
/* select the k index used by the next solve (negf%HS(iK)) and its weight */
void negf_set_kpoint(int *handler, int kpoint);

/* solves (libnegf_api.f90:689-717, 1205-1236) */
void negf_solve_landauer(int *handler);                  /* compute_current            */
void negf_calculate_dephasing_transmission(int *handler);
void negf_current(int *handler, double *current, char *unitsOfH, char *unitsOfJ);

/* elastic dephasing (libnegf_api.f90:1289-1310, 1440-1457) */
void negf_set_elph_dephasing(int *handler, double *coupling, int coupling_size, int scba_niter);
void negf_set_scba_tolerances(int *handler, double elastic_tol, double inelastic_tol);

/* results (libnegf_api.f90:802-935); associate_* return library-owned
 * column-major arrays valid until the next solve or destruct. */
void negf_get_energies(int *handler, int *npoints, double *re_en, double *im_en, int copy);
void negf_get_currents(int *handler, int *npairs, double *currents, int copy);
void negf_associate_transmission(int *handler, int tr_shape[2], double **tr_pointer);
void negf_associate_energy_current(int *handler, int currents_shape[2], double **currents_pointer);

/* Files (libnegf_api.f90:289-309, 589-628, 1032-1043): negf_read_hs reads H
 * (target 0) or S (target 1) from the formatted PETSc ASCII pair of files
 * (real part, imaginary part; read_H, input_output.F90:41-264, full matrix);
 * negf_write_tunneling_and_dos writes tunneling_<kpoint>_000000.dat and
 * localDOS_<kpoint>_000000.dat (Fortran E17.8 columns) into the path set by
 * negf_set_output (default "./"). */
void negf_read_hs(int *handler, const char *real_path, const char *imag_path, int target_matrix);
void negf_set_output(int *handler, const char *out_path);
void negf_write_tunneling_and_dos(int *handler);

/* Scratch folder of the surface-function disk cache (negf_set_scratch,
 * libnegf_api.f90:314-334; default "./GS/", which the caller creates).  With
 * lnparams.sgfcache = 0 (disk) and readold_t_sgfs 0 (load, else compute and
 * save) or 2 (compute and save), the Landauer, LDOS and dephasing solves read
 * and write g_s of every decimated contact and energy point there, one file
 * Mat<u|d><contact:2>_00_<kpoint>_<point>.dat each, in the reference's
 * format (ln_cache.F90:458-594, outmatrix.F90:187-245: Fortran unformatted
 * sequential records of int32 i, int32 j, complex(dp) value, for the entries
 * that are not zero); sgfcache = 1 keeps them in device memory, 2 never
 * caches.  The density paths use the device-memory cache. */
void negf_set_scratch(int *handler, const char *scratch_path);

/* Density matrix (negf_solve_density_dft, negf_get_dm, negf_set_outer;
 * libnegf_api.f90:780-791, 836-861, 1052-1065): compute_density_dft
 * (libnegf.F90:1664-1679) without inelastic terms -- the contour integral
 * (Np_n(1) + Np_n(2) points and n_poles poles from lnparams.ec up to the
 * reference contact's mu, chosen by min_or_max) and, when np_real > 0, the
 * real-axis integral of G^n between the contacts' Fermi levels.  rho is
 * returned as 1-based CSR (nrow = the size of H) on the pattern: diagonal,
 * S pattern, the device/contact coupling pattern (outer blocks), and their
 * transposes.  DorE = 'E' gives the energy-weighted matrix. */
void negf_solve_density_dft(int *handler);
void negf_get_dm(int *handler, int *nnz, int *nrow, int *rowpnt, int *colind, double *re_nzval,
                 double *im_nzval, int copy);
void negf_set_outer(int *handler, int *outer);
/* Electron (particle 1) or hole (-1) density of compute_density_efa
 * (negf_density_efa, libnegf_api.f90:724-749; libnegf.F90:1690-1791):
 * contour from Ec - DeltaEc (holes: Ev + DeltaEv) with Np_n (Np_p) points
 * and n_poles poles around mu_n (mu_p) of contact 1, then np_real trapezoid
 * points on the real axis (integrations.F90:281-592, 1346-1462; holes
 * integrate A - G^n).  density[k] = Re rho(k, k) for k < ndofs; rho stays
 * readable with negf_get_dm.  contacts = true (the bulk contact density) is
 * not provided and reports an error. */
void negf_density_efa(int *handler, int *ndofs, double *density, int *particle, bool *contacts);
/* Quasi-equilibrium electron (particle 1) or hole (-1) density per orbital
 * (negf_density_quasi_equilibrium, libnegf_api.f90:753-775; quasiEq_int_n/p,
 * integrations.F90:3002-3258) for band-edge profiles Ec/Ev and quasi-Fermi
 * levels mu_n/mu_p of ndofs orbitals: a contour over every orbital's window
 * (Np(2) horizontal and Np(1) vertical points) and the poles of each
 * orbital's own level; density[k] = Re i (t_k - conj t_k). */
void negf_density_quasi_equilibrium(int *handler, int *ndofs, double *density, int *particle, double *Ec,
                                    double *Ev, double *mu_n, double *mu_p);

/* Local DOS alongside the transmission (negf_init_ldos / negf_set_ldos_intervals /
 * negf_set_ldos_indexes / negf_associate_ldos, libnegf_api.f90:941-1021):
 * once projections are set, negf_solve_landauer also fills ldos_mat
 * (Npoints x nldos, column-major, times kwght) from
 * calculate_transmissions_and_dos at E + i dos_delta (integrations.F90:1758-1766). */
void negf_init_ldos(int *handler, int nldos);
void negf_set_ldos_intervals(int *handler, int nldos, int *istart, int *iend);
void negf_set_ldos_indexes(int *handler, int ildos, int idx_size, int *idx);
void negf_associate_ldos(int *handler, int ldos_shape[2], double **ldos_pointer);
/* compute_ldos (libnegf.F90:1893-1901, Fortran API only in the reference):
 * ldos_mat(i, l) = sum over projection l's orbitals of -g_spin Im diag(G^r S) / pi
 * at E_i + i dos_delta on the transmission grid (ldos_int, integrations.F90:212-258;
 * G^r with its device/contact blocks, no kwght).  Fails with elastic
 * interactions and scba_niter > 0. */
void negf_compute_ldos(int *handler);

/* extensions of this implementation */
void negf_set_precision(int *handler, int precision); /* hipnegf_precision (FP64 default) */
void negf_set_batch(int *handler, int max_batch);     /* 0: automatic                  */
/* tun_proj (set_tun_indexes, libnegf.F90:1044-1058, which the reference's C API
 * does not export): 1-based device orbitals whose rows enter the transmission
 * traces; n < 0 removes it (all rows). */
void negf_set_tun_indexes(int *handler, int n, int *idx);
int negf_get_status(int *handler);                    /* 0 ok, else last failure       */
const char *negf_get_last_error(int *handler);
/* per-energy status bits and decimation iteration counts of the last solve */
void negf_get_item_status(int *handler, int *npoints, int *status, int *ncyc, int copy);

#ifdef __cplusplus
}
#endif

#endif /* HIPNEGF_API_H */
