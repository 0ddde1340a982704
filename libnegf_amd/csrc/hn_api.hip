// hn_api.hip -- libNEGF public C API subset (include/hipnegf_api.h) on the
// batched MI355X engine.
//
// Host logic mirrors the reference's drivers:
//   negf_init / set_defaults            lib_param.F90:602-669, libnegf.F90:148-181
//   set_H / set_S / set_S_id            libnegf.F90:303-434 (0/1-based CSR)
//   init_structure / create_TStruct     libnegf.F90:555-590, ln_structure.F90:80-150
//   extract_cont + zcsr2blk_sod         ln_extract.F90:41-76, sparsekit_drv.F90:3409-3436
//   compute_current                     libnegf.F90:1904-1963
//   tunneling_int_def / tunneling_and_dos  integrations.F90:1613-1783
//   init_ldos / set_ldos_* / associate_ldos libnegf.F90:971-1041, 1416-1426
//   meir_wingreen (fixed occupations)   integrations.F90:1804-1966
//   electron_current / integrate_el     integrations.F90:2427-2457, 2688-2767
// Every (E) point of a solve is sent to the engine in one batched call.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/hipnegf.h"
#include "../../include/hipnegf_api.h"
#include "hn_engine.h"

namespace {

struct Csr {
  int n = 0;
  std::vector<double2> v;
  std::vector<int> col, ptr;  // 0-based
  bool present() const { return n > 0; }
};

struct Session {
  lnparams p{};
  int ncont = 0, npl = 0;
  bool structure_set = false;
  std::vector<int> surfstart, surfend, contend, plend, cblk;  // 1-based
  std::vector<Csr> H, S;
  std::vector<char> s_id;
  int cur = 0;  // k slot of the working H/S pointers (negf%H => HS(iK)%H)
  bool deph = false;
  std::vector<double> coupling;
  int scba_niter = 0;
  double tol_el = 1e-7, tol_inel = 1e-7;
  int precision = 0, max_batch = 0;
  // results
  int npoints = 0, npairs = 0;
  std::vector<double> en, en_im, tunn, curr_mat, currents;
  std::vector<int> item_status, item_ncyc;
  int status = 0;
  std::string err;
  std::unique_ptr<hn::Engine> eng;
  // device copy of the working H/S, rebuilt only after H, S or the partition change
  std::unique_ptr<hn::DevStruct> dev;
  long long gen = 0, dev_gen = -1;
  // the same structure padded to one block size, for the SCBA path
  std::unique_ptr<hn::DevStruct> dev_pad;
  long long dev_pad_gen = -1;
  bool sgf_cache_on = false;
  // LDOS projections (negf%dos_proj / ndos_proj, libnegf.F90:971-1041), their
  // result ldos_mat (npoints x ndos, column-major) and tun_proj (:1044-1058)
  int ndos = 0;
  std::vector<std::vector<int>> dos_proj;
  std::vector<double> ldos;
  int ldos_cols = 0;
  bool tun_set = false;
  std::vector<int> tun_proj;
  std::string out_path = "./";  // negf%out_path (set_outpath appends '/')
  std::string scratch = "./GS/";  // negf%scratch_path (lib_param.F90:607; set_scratch appends '/')
  // density matrix (negf%rho, CSR 0-based here) and negf%outer_blocks
  int outer = 2;
  int rho_n = 0;
  std::vector<int> rho_ptr, rho_col;
  std::vector<double2> rho_val;
};

Session *sess(int *handler) {
  Session *s = nullptr;
  std::memcpy(&s, handler, sizeof(s));
  return s;
}

template <typename F> void api(int *handler, F &&f) {
  Session *s = sess(handler);
  if (!s) return;
  try {
    s->status = 0;
    s->err.clear();
    f(*s);
  } catch (const std::exception &e) {
    s->status = 1;
    s->err = e.what();
    if (s->p.verbose > 0) std::fprintf(stderr, "hipnegf: %s\n", e.what());
  }
}

// set_defaults (lib_param.F90:602-669) and init_contacts defaults (libnegf.F90:612-646)
void set_defaults(lnparams &p) {
  std::memset(&p, 0, sizeof(p));
  p.verbose = 10;
  p.readold_dm_sgfs = 1;
  p.readold_t_sgfs = 1;
  p.sgfcache = 0;
  p.spin = 1;
  p.kpoint = 1;
  p.g_spin = 2.0;
  p.delta = 1.0e-4;
  p.deltaModel = 1;
  p.wmax = 0.009;
  p.dos_delta = 0.0;
  p.eneconv = 1.0;
  p.kwght = 1.0;
  p.scba_inelastic_tol = 1.0e-7;
  p.scba_elastic_tol = 1.0e-7;
  p.np_n[0] = p.np_n[1] = 20;
  p.np_p[0] = p.np_p[1] = 20;
  p.n_kt = 10;
  p.n_poles = 3;
  p.ni[0] = 1;
  p.nf[0] = 2;
  p.dore[0] = 'D';
  p.min_or_max = 1;
}

// set_H: 0-based when rowpnt(1) == 0 (libnegf.F90:324-341)
Csr make_csr(int nrow, const double2 *A, const int *JA, const int *IA) {
  Csr c;
  c.n = nrow;
  const int base = (IA[0] == 0) ? 0 : 1;
  const int nnz = IA[nrow] - IA[0];
  c.ptr.resize(nrow + 1);
  for (int i = 0; i <= nrow; ++i) c.ptr[i] = IA[i] - IA[0];
  c.col.resize(nnz);
  c.v.resize(nnz);
  for (int i = 0; i < nnz; ++i) {
    c.col[i] = JA[i] - base;
    c.v[i] = A[i];
  }
  return c;
}

// dense column-major extract of rows [r1, r2], cols [c1, c2] (1-based inclusive)
std::vector<double2> extract(const Csr &A, int r1, int r2, int c1, int c2) {
  const int m = r2 - r1 + 1, n = c2 - c1 + 1;
  std::vector<double2> out((size_t)m * n, make_double2(0.0, 0.0));
  if (m <= 0 || n <= 0) return out;
  if (r2 > A.n || c2 > A.n) throw std::runtime_error("extract: block outside the matrix");
  for (int i = r1 - 1; i < r2; ++i)
    for (int k = A.ptr[i]; k < A.ptr[i + 1]; ++k) {
      const int j = A.col[k];
      if (j >= c1 - 1 && j < c2) {
        double2 &d = out[(i - r1 + 1) + (size_t)(j - c1 + 1) * m];
        d.x += A.v[k].x;  // duplicates are summed (CSR sum semantics)
        d.y += A.v[k].y;
      }
    }
  return out;
}

std::vector<double2> eye(int n) {
  std::vector<double2> out((size_t)n * n, make_double2(0.0, 0.0));
  for (int i = 0; i < n; ++i) out[i + (size_t)i * n] = make_double2(1.0, 0.0);
  return out;
}

hn::Engine &engine(Session &s) {
  if (!s.eng) {
    int dev = 0;
    HN_CHECK(hipGetDevice(&dev));
    s.eng.reset(new hn::Engine(dev));
  }
  return *s.eng;
}

// Device structure of the working H/S: the tridiagonal PL blocks
// (zcsr2blk_sod) and the contact blocks (extract_cont, ln_extract.F90:41-76):
// the surface block (surfstart..surfend, may be empty) and the two PLs after
// it; HMC/SMC span the surface and the first PL.
// With pad > 0 every PL and contact layer is stored as a pad x pad block:
// the physical block in the top-left corner, and each padding orbital
// decoupled with H = -1, S = 0 on its diagonal, so that E S - H = 1 there and
// the padding adds exact zeros to every product the physical entries see.
// This serves the SCBA path, whose sweep runs on blocks of one size.
std::unique_ptr<hn::DevStruct> build_struct(Session &s, int pad) {
  if (!s.structure_set) throw std::runtime_error("structure not initialised (negf_init_structure)");
  if (s.cur >= (int)s.H.size() || !s.H[s.cur].present()) throw std::runtime_error("H not set");
  const Csr &H = s.H[s.cur];
  const bool sid = s.cur >= (int)s.S.size() || s.s_id[s.cur] || !s.S[s.cur].present();
  const Csr *S = sid ? nullptr : &s.S[s.cur];
  const int nbl = s.npl;
  std::vector<int> pls(nbl), ple(nbl), bsz(nbl);
  for (int i = 0; i < nbl; ++i) {
    pls[i] = (i == 0) ? 1 : s.plend[i - 1] + 1;
    ple[i] = s.plend[i];
    bsz[i] = ple[i] - pls[i] + 1;
  }
  std::vector<int> nc(s.ncont), ns(s.ncont);
  for (int c = 0; c < s.ncont; ++c) {
    const int surfdim = (s.surfend[c] + 1) - s.surfstart[c];  // mat_C_start - mat_B_start
    if (surfdim < 0) throw std::runtime_error("contact surface end before its start");
    const int ncdim = s.contend[c] - s.surfend[c];           // the two PLs after the surface
    if (ncdim <= 0 || ncdim % 2) throw std::runtime_error("contact must hold 2 principal layers");
    nc[c] = ncdim / 2;
    ns[c] = surfdim;
    if (s.cblk[c] < 1 || s.cblk[c] > nbl) throw std::runtime_error("cblk out of range");
  }
  if (pad > 0)
    for (int c = 0; c < s.ncont; ++c)
      if (ns[c]) throw std::runtime_error("SCBA path requires equal PL and contact sizes with a surface block");
  const bool padS = pad > 0 && sid;  // padding needs an explicit S
  const int smode = (sid && !padS) ? 0 : (hn::S_DIAG | hn::S_OFFDIAG | hn::S_CONTACT);
  std::vector<int> bsz_d = bsz, nc_d = nc;
  if (pad > 0) {
    bsz_d.assign(nbl, pad);
    nc_d.assign(s.ncont, pad);
  }
  std::unique_ptr<hn::DevStruct> st(new hn::DevStruct(nbl, bsz_d.data(), s.ncont, s.cblk.data(),
                                                      nc_d.data(), smode, 1, ns.data()));
  // kinds whose padding diagonal is -1 (H) or 0 (S); off-diagonal blocks pad with zeros
  auto diag_kind = [](int kind) { return kind == hn::BK_HD || kind == hn::BK_H00; };
  auto put = [&](int kind, int idx, const std::vector<double2> &h0) {
    const long long n = st->block_elems(kind, idx);
    std::vector<double2> h = h0;
    if (pad > 0 && (long long)h0.size() != n) {  // top-left corner of a pad x pad block
      const int r = (int)std::lround(std::sqrt((double)h0.size()));
      const bool sq = (long long)r * r == (long long)h0.size();
      int rows = r, cols = r;
      if (!sq || (kind == hn::BK_HU || kind == hn::BK_HL || kind == hn::BK_SU || kind == hn::BK_SL ||
                  kind == hn::BK_HMC || kind == hn::BK_SMC))
        throw std::runtime_error("padded block of unknown shape");
      h.assign((size_t)pad * pad, make_double2(0.0, 0.0));
      for (int j = 0; j < cols; ++j)
        for (int i = 0; i < rows; ++i) h[i + (size_t)j * pad] = h0[i + (size_t)j * rows];
      if (diag_kind(kind))
        for (int i = rows; i < pad; ++i) h[i + (size_t)i * pad] = make_double2(-1.0, 0.0);
    }
    if ((long long)h.size() != n) throw std::runtime_error("block size mismatch");
    HN_CHECK(hipMemcpy(st->pool + st->off[kind][idx], h.data(), n * sizeof(double2),
                       hipMemcpyHostToDevice));
  };
  // rectangular blocks: rows x cols in the top-left corner of pad x pad
  auto putr = [&](int kind, int idx, const std::vector<double2> &h0, int rows, int cols) {
    if (pad <= 0) return put(kind, idx, h0);
    std::vector<double2> h((size_t)pad * pad, make_double2(0.0, 0.0));
    for (int j = 0; j < cols; ++j)
      for (int i = 0; i < rows; ++i) h[i + (size_t)j * pad] = h0[i + (size_t)j * rows];
    const long long n = st->block_elems(kind, idx);
    if ((long long)h.size() != n) throw std::runtime_error("block size mismatch");
    HN_CHECK(hipMemcpy(st->pool + st->off[kind][idx], h.data(), n * sizeof(double2),
                       hipMemcpyHostToDevice));
  };
  auto eye = [](int n) {
    std::vector<double2> e((size_t)n * n, make_double2(0.0, 0.0));
    for (int i = 0; i < n; ++i) e[i + (size_t)i * n] = make_double2(1.0, 0.0);
    return e;
  };
  for (int i = 0; i < nbl; ++i) {
    put(hn::BK_HD, i, extract(H, pls[i], ple[i], pls[i], ple[i]));
    if (S) put(hn::BK_SD, i, extract(*S, pls[i], ple[i], pls[i], ple[i]));
    else if (padS) put(hn::BK_SD, i, eye(bsz[i]));
  }
  for (int i = 0; i + 1 < nbl; ++i) {
    putr(hn::BK_HU, i, extract(H, pls[i], ple[i], pls[i + 1], ple[i + 1]), bsz[i], bsz[i + 1]);
    putr(hn::BK_HL, i, extract(H, pls[i + 1], ple[i + 1], pls[i], ple[i]), bsz[i + 1], bsz[i]);
    if (S) {
      putr(hn::BK_SU, i, extract(*S, pls[i], ple[i], pls[i + 1], ple[i + 1]), bsz[i], bsz[i + 1]);
      putr(hn::BK_SL, i, extract(*S, pls[i + 1], ple[i + 1], pls[i], ple[i]), bsz[i + 1], bsz[i]);
    } else if (padS) {
      putr(hn::BK_SU, i, std::vector<double2>((size_t)bsz[i] * bsz[i + 1]), bsz[i], bsz[i + 1]);
      putr(hn::BK_SL, i, std::vector<double2>((size_t)bsz[i] * bsz[i + 1]), bsz[i + 1], bsz[i]);
    }
  }
  for (int c = 0; c < s.ncont; ++c) {
    const int c0 = s.surfstart[c], q = ns[c], c1 = c0 + q, n = nc[c];  // surface at c0, PL1 at c1
    const int b = s.cblk[c] - 1;
    put(hn::BK_H00, c, extract(H, c1, c1 + n - 1, c1, c1 + n - 1));
    putr(hn::BK_H01, c, extract(H, c1, c1 + n - 1, c1 + n, c1 + 2 * n - 1), n, n);
    putr(hn::BK_HMC, c, extract(H, pls[b], ple[b], c0, c1 + n - 1), bsz[b], q + n);
    if (q > 0) {
      put(hn::BK_HSS, c, extract(H, c0, c1 - 1, c0, c1 - 1));
      put(hn::BK_HS1, c, extract(H, c0, c1 - 1, c1, c1 + n - 1));
      put(hn::BK_H1S, c, extract(H, c1, c1 + n - 1, c0, c1 - 1));
    }
    if (padS) {
      putr(hn::BK_S00, c, eye(n), n, n);
      putr(hn::BK_S01, c, std::vector<double2>((size_t)n * n), n, n);
      putr(hn::BK_SMC, c, std::vector<double2>((size_t)bsz[b] * n), bsz[b], n);
    }
    if (S) {
      putr(hn::BK_S00, c, extract(*S, c1, c1 + n - 1, c1, c1 + n - 1), n, n);
      putr(hn::BK_S01, c, extract(*S, c1, c1 + n - 1, c1 + n, c1 + 2 * n - 1), n, n);
      putr(hn::BK_SMC, c, extract(*S, pls[b], ple[b], c0, c1 + n - 1), bsz[b], q + n);
      if (q > 0) {
        put(hn::BK_SSS, c, extract(*S, c0, c1 - 1, c0, c1 - 1));
        put(hn::BK_SS1, c, extract(*S, c0, c1 - 1, c1, c1 + n - 1));
        put(hn::BK_S1S, c, extract(*S, c1, c1 + n - 1, c0, c1 - 1));
      }
    }
  }
  return st;
}

// fermi_f, distributions.F90:46-74
double fermi(double E, double Ef, double kT) {
  if (kT == 0.0) return E > Ef ? 0.0 : 1.0;
  if (std::fabs((E - Ef) / kT) > 30) return E > Ef ? std::exp(-(E - Ef) / kT) : 1.0;
  return 1.0 / (std::exp((E - Ef) / kT) + 1.0);
}

// integrate_el, integrations.F90:2688-2767 (Kb in Hartree/K, ln_constants.F90:36)
double integrate_el(const double *T, double mu1, double mu2, double kT1, double kT2, double emin,
                    double emax, double estep) {
  const double Kb = 3.166811389002312e-6;
  double curr = 0.0, destep = 1.0e10;
  int N = 0;
  const int Np = (int)std::lround((emax - emin) / estep);
  double kbT1 = kT1 < Kb ? Kb : kT1, kbT2 = kT2 < Kb ? Kb : kT2;
  bool swapped = false;
  if (mu2 < mu1) {
    swapped = true;
    std::swap(mu1, mu2);
    std::swap(kbT1, kbT2);
  }
  for (int i = 0; i < Np; ++i) {
    const double E1 = emin + estep * i, TT1 = T[i];
    const double E2 = emin + estep * (i + 1), TT2 = T[i + 1];
    while (destep >= (kbT1 + kbT2) / 20.0) {
      N = N + 1;
      destep = (E2 - E1) / N;
    }
    for (int i1 = 0; i1 < N; ++i1) {
      const double E3 = E1 + (E2 - E1) * i1 / N;
      const double E4 = E3 + (E2 - E1) / N;
      const double TT3 = (TT2 - TT1) * i1 / N + TT1;
      const double TT4 = TT3 + (TT2 - TT1) / N;
      const double c1 = (fermi(E3, mu2, kbT2) - fermi(E3, mu1, kbT1)) * TT3;
      const double c2 = (fermi(E4, mu2, kbT2) - fermi(E4, mu1, kbT1)) * TT4;
      curr = curr + (c1 + c2) * (E4 - E3) / 2.0;
    }
  }
  return swapped ? -curr : curr;
}

std::vector<int> pairs_of(const int *v) {
  std::vector<int> out;
  for (int i = 0; i < MAXNLAYERS && v[i] != 0; ++i) out.push_back(v[i]);
  return out;
}

// The device structure of the working H/S (built once per change of H, S or
// the partition); fictitious-contact parameters are applied on every call.
hn::DevStruct &device_struct(Session &s) {
  if (!s.dev || s.dev_gen != s.gen) {
    s.dev = build_struct(s, 0);
    s.dev_gen = s.gen;
  }
  hn::DevStruct &d = *s.dev;
  if (s.tun_set) {
    // an empty tun_proj selects no row (get_tun_mask with a zero-size list);
    // index 0 marks no orbital
    const int zero = 0;
    if (s.tun_proj.empty()) d.set_tun_proj(1, &zero);
    else d.set_tun_proj((int)s.tun_proj.size(), s.tun_proj.data());
  } else if (d.tun_mask) {
    d.set_tun_proj(0, nullptr);
  }
  for (int c = 0; c < d.ncont; ++c) {
    const int f = s.p.fictcont[c] ? 1 : 0;
    if (f != d.fict[c] || (f && d.contact_dos[c] != s.p.contact_dos[c])) {
      d.fict[c] = f;
      d.contact_dos[c] = s.p.contact_dos[c];
      d.touched();
    }
  }
  return d;
}

// The structure the SCBA path runs on: the device copy itself when every PL
// and contact layer has one size, else a copy padded to the largest (see
// build_struct); pad returns that size (0: not padded).
hn::DevStruct &scba_struct(Session &s, int &pad) {
  hn::DevStruct &d = device_struct(s);
  pad = 0;
  if (d.uniform()) return d;
  for (int b : d.bsz) pad = std::max(pad, b);
  for (int c = 0; c < d.ncont; ++c) pad = std::max(pad, d.nc[c]);
  if (!s.dev_pad || s.dev_pad_gen != s.gen) {
    s.dev_pad = build_struct(s, pad);
    s.dev_pad_gen = s.gen;
  }
  hn::DevStruct &q = *s.dev_pad;
  for (int c = 0; c < q.ncont; ++c)
    if (q.fict[c] != d.fict[c] || q.contact_dos[c] != d.contact_dos[c]) {
      q.fict[c] = d.fict[c];
      q.contact_dos[c] = d.contact_dos[c];
      q.touched();
    }
  return q;
}

hn::SolveOpts opts_of(Session &s) {
  hn::SolveOpts o;
  // readOldT_SGFs (lib_param.F90): 0 load cached g_s (else compute and save),
  // 1 compute, 2 compute and save.  SGFcache 1 keeps them in device memory,
  // 0 in files of the scratch folder (staged through the device cache, see
  // DiskSgf), 2 nowhere (TMatrixCacheDummy: never cached, so always computed)
  const int r = s.p.readold_t_sgfs;
  o.sgf_mode = s.p.sgfcache == 2 ? 0 : (r == 0) ? 1 : (r == 2) ? 2 : 0;
  if (o.sgf_mode && !s.sgf_cache_on) {
    engine(s).set_sgf_cache(2ull << 30);
    s.sgf_cache_on = true;
  }
  o.delta = s.p.delta;
  o.precision = s.precision;
  o.max_batch = s.max_batch;
  return o;
}

// ---------------------------------------------------------------------------
// Disk cache of surface Green's functions (SGFcache = 0; ln_cache.F90
// TMatrixCacheDisk :458-594, outmat_c / inmat_c outmatrix.F90:187-245): one
// file per (spin, contact, k-point, energy point) in the scratch folder,
// named Mat<u|d><contact, 2 digits>_00_<k-point>_<point>.dat.  A file is a
// Fortran unformatted sequential file holding one record per entry of g_s
// that is not zero (outmat_c's threshold EPS is never set by the library, so
// it is 0), in column order: int32 i, int32 j (1-based), complex(dp) g(i,j).
// As gfortran writes it, each record is framed by its byte length (int32, 24)
// before and after.  Reading starts from zeros and sets the entries listed.
// ---------------------------------------------------------------------------
std::string label_str(int v) {  // get_string_label, ln_cache.F90:669-679
  if (v > 9999999) throw std::runtime_error("ERROR: label too large (> 9999999)");
  const int w = v > 999999 ? 7 : v > 99999 ? 6 : v > 9999 ? 5 : v > 999 ? 4 : v > 99 ? 3 : 2;
  char b[16];
  if (v < 0) std::snprintf(b, sizeof(b), "%s", "**");  // a sign does not fit i2.2: Fortran prints **
  else std::snprintf(b, sizeof(b), "%0*d", w, v);
  return b;
}

std::string sgf_filename(const Session &s, int contact, int point) {  // disk_indices_to_filename
  if (s.p.spin != 1 && s.p.spin != 2) throw std::runtime_error("surface-function files: spin must be 1 or 2");
  char row[8];
  std::snprintf(row, sizeof(row), "%02d", contact);
  return s.scratch + "Mat" + (s.p.spin == 1 ? "u" : "d") + row + "_00_" + label_str(s.p.kpoint) + "_" +
         label_str(point) + ".dat";
}

void sgf_write_file(const std::string &fn, const double2 *g, int n) {
  FILE *f = std::fopen(fn.c_str(), "wb");
  if (!f) throw std::runtime_error("cannot open " + fn + " (does the scratch folder exist?)");
  std::vector<char> buf;
  buf.reserve((size_t)n * n * 32);
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) {
      const double2 v = g[(size_t)j * n + i];
      if (!(std::fabs(v.x) > 0.0 || std::fabs(v.y) > 0.0)) continue;
      const int32_t len = 24, ii = i + 1, jj = j + 1;
      char r[32];
      std::memcpy(r, &len, 4);
      std::memcpy(r + 4, &ii, 4);
      std::memcpy(r + 8, &jj, 4);
      std::memcpy(r + 12, &v.x, 8);
      std::memcpy(r + 20, &v.y, 8);
      std::memcpy(r + 28, &len, 4);
      buf.insert(buf.end(), r, r + 32);
    }
  const bool ok = std::fwrite(buf.data(), 1, buf.size(), f) == buf.size();
  if (std::fclose(f) != 0 || !ok) throw std::runtime_error("cannot write " + fn);
}

bool sgf_read_file(const std::string &fn, std::vector<double2> &g, int n) {
  FILE *f = std::fopen(fn.c_str(), "rb");
  if (!f) return false;
  g.assign((size_t)n * n, double2{0.0, 0.0});
  char r[32];
  size_t got;
  while ((got = std::fread(r, 1, 32, f)) > 0) {
    int32_t l0, l1, i, j;
    double2 v;
    std::memcpy(&l0, r, 4);
    std::memcpy(&l1, r + 28, 4);
    std::memcpy(&i, r + 4, 4);
    std::memcpy(&j, r + 8, 4);
    std::memcpy(&v.x, r + 12, 8);
    std::memcpy(&v.y, r + 20, 8);
    if (got != 32 || l0 != 24 || l1 != 24 || i < 1 || i > n || j < 1 || j > n) {
      std::fclose(f);
      throw std::runtime_error("malformed surface-function file " + fn);
    }
    g[(size_t)(j - 1) * n + (i - 1)] = v;
  }
  std::fclose(f);
  return true;
}

// Wraps one real-axis solve of npoints energies (pt_path = point index + 1,
// integrations.F90:1637): with SGFcache 0 and readOldT_SGFs 0 the files that
// exist are loaded into the device cache before the solve (a point whose
// files are all there is not recomputed, ncyc 0); after the solve g_s is
// written for every point that had no file (readOldT_SGFs 0) or for every
// point (readOldT_SGFs 2), as surface_green does (contselfenergy.F90:161-259).
// On a structure padded for the SCBA path the files hold the physical
// top-left block; the padding orbitals' g_s is exactly 1 on the diagonal
// (E S - H = 1 there, decoupled), which a load restores.
struct DiskSgf {
  Session &s;
  hn::DevStruct &st;
  hn::Engine &eng;
  hn::SolveOpts o;
  bool on = false;
  std::vector<char> save;  // (point, contact) written after the solve
  int phys(int c) const { return s.dev ? s.dev->ngs(c) : st.ngs(c); }
  DiskSgf(Session &s_, hn::DevStruct &st_, hn::Engine &eng_, const hn::SolveOpts &o_)
      : s(s_), st(st_), eng(eng_), o(o_) {
    on = s.p.sgfcache == 0 && o.sgf_mode != 0;
    if (!on) return;
    const int n = s.npoints, nc = st.ncont;
    long long ent = 0;
    int live = 0;
    for (int c = 0; c < nc; ++c)
      if (!st.fict[c]) {
        ent = std::max(ent, (long long)st.ngs(c) * st.ngs(c));
        ++live;
      }
    if (!live) { on = false; return; }
    // every entry of this solve stays resident until it is written
    const size_t need = (size_t)ent * sizeof(double2) * (size_t)n * live + ((size_t)16 << 20);
    if (need > ((size_t)64 << 30))
      throw std::runtime_error("surface-function files: more than 64 GiB of g_s in one solve");
    if (eng.sgf_cache_bytes() < need) {
      eng.set_sgf_cache(need);
      s.sgf_cache_on = true;
    }
    save.assign((size_t)n * nc, 0);
    std::vector<double2> g;
    for (int i = 0; i < n; ++i)
      for (int c = 0; c < nc; ++c) {
        if (st.fict[c]) continue;
        const int np = phys(c), ng = st.ngs(c);
        const bool have = o.sgf_mode == 1 && sgf_read_file(sgf_filename(s, c + 1, i + 1), g, np);
        if (have) {
          if (ng != np) {  // expand into the padded block
            std::vector<double2> e((size_t)ng * ng, make_double2(0.0, 0.0));
            for (int j = 0; j < ng; ++j)
              for (int r = 0; r < ng; ++r)
                e[r + (size_t)j * ng] = (r < np && j < np) ? g[r + (size_t)j * np]
                                        : make_double2(r == j ? 1.0 : 0.0, 0.0);
            g.swap(e);
          }
          eng.sgf_insert(st, o, 0, s.en[i], c, g.data(), 0);
        } else {
          save[(size_t)i * nc + c] = 1;
        }
      }
  }
  void finish() {
    if (!on) return;
    const int n = s.npoints, nc = st.ncont;
    std::vector<double2> g;
    for (int i = 0; i < n; ++i)
      for (int c = 0; c < nc; ++c) {
        if (!save[(size_t)i * nc + c]) continue;
        const int np = phys(c), ng = st.ngs(c);
        g.resize((size_t)ng * ng);
        int stat = 0;
        if (!eng.sgf_fetch(st, o, 0, s.en[i], c, g.data(), &stat)) continue;
        if (ng != np)  // the physical top-left block, column by column
          for (int j = 0; j < np; ++j)
            for (int r = 0; r < np; ++r) g[r + (size_t)j * np] = g[r + (size_t)j * ng];
        sgf_write_file(sgf_filename(s, c + 1, i + 1), g.data(), np);
      }
  }
};

// electron_current, integrations.F90:2427-2457
void electron_current(Session &s) {
  const std::vector<int> ni = pairs_of(s.p.ni), nf = pairs_of(s.p.nf);
  s.currents.assign(s.npairs, 0.0);
  if (s.ncont < 2 || s.tunn.empty()) return;
  for (int ii = 0; ii < s.npairs; ++ii) {
    const int a = ni[ii] - 1, b = nf[ii] - 1;
    s.currents[ii] = integrate_el(s.tunn.data() + (size_t)ii * s.npoints, s.p.mu[a], s.p.mu[b],
                                  s.p.kbt_t[a], s.p.kbt_t[b], s.p.emin, s.p.emax, s.p.estep) *
                     s.p.g_spin;
  }
}

void energy_grid(Session &s) {
  // tunneling_int_def, integrations.F90:1613-1668 (one process: no padding)
  s.en_im.clear();
  s.npoints = (int)std::lround((s.p.emax - s.p.emin) / s.p.estep) + 1;
  s.en.resize(s.npoints);
  for (int i = 0; i < s.npoints; ++i) s.en[i] = s.p.emin + s.p.estep * i;
}

struct DevBuf {
  void *p = nullptr;
  explicit DevBuf(size_t bytes) { HN_CHECK(hipMalloc(&p, bytes ? bytes : 1)); }
  ~DevBuf() { if (p) (void)hipFree(p); }
  template <typename T> T *as() { return (T *)p; }
};

// tunneling_and_dos with LDOS (integrations.F90:1724-1766): T and ldos_mat
// rows from calculate_transmissions_and_dos at E + i dos_delta, both times kwght.
void compute_landauer_dos(Session &s, const std::vector<int> &ni, const std::vector<int> &nf) {
  hn::DevStruct *st = &device_struct(s);
  hn::Engine &eng = engine(s);
  const int n = s.npoints, ndev = st->ndev();
  // dos_proj(1)%indexes(1) == 0 with one projection per orbital: every orbital
  // (iterative.F90:1930-1931)
  const bool full = s.ndos == ndev && !s.dos_proj[0].empty() && s.dos_proj[0][0] == 0;
  std::vector<int> ptr(1, 0), idx;
  if (!full)
    for (int l = 0; l < s.ndos; ++l) {
      for (int i2 : s.dos_proj[l]) {
        if (i2 < 1) throw std::runtime_error("LDOS orbital index < 1");
        idx.push_back(i2);
      }
      ptr.push_back((int)idx.size());
    }
  const int nout = full ? ndev : s.ndos;
  const int nc = std::max(s.ncont, 1);
  DevBuf dE(sizeof(double) * n), dsl(sizeof(int) * n), dN(sizeof(int) * n * nc), dS(sizeof(int) * n),
      dTp(sizeof(double) * n * s.npairs), dL(sizeof(double) * (size_t)n * nout),
      dP(sizeof(int) * (ptr.size() + idx.size() + 1));
  HN_CHECK(hipMemcpy(dE.p, s.en.data(), sizeof(double) * n, hipMemcpyHostToDevice));
  HN_CHECK(hipMemset(dsl.p, 0, sizeof(int) * n));
  hn::DosOpts d;
  d.dos_delta = s.p.dos_delta;
  d.full = full ? 1 : 0;
  if (!full) {
    HN_CHECK(hipMemcpy(dP.p, ptr.data(), sizeof(int) * ptr.size(), hipMemcpyHostToDevice));
    if (!idx.empty())
      HN_CHECK(hipMemcpy(dP.as<int>() + ptr.size(), idx.data(), sizeof(int) * idx.size(),
                         hipMemcpyHostToDevice));
    d.nldos = s.ndos;
    d.ptr = dP.as<int>();
    d.idx = dP.as<int>() + ptr.size();
  }
  hn::SolveOpts o = opts_of(s);
  DiskSgf disk(s, *st, eng, o);
  eng.transmission_dos(*st, n, dE.as<double>(), dsl.as<int>(), o, s.npairs, ni.data(), nf.data(), d,
                       dTp.as<double>(), dL.as<double>(), dN.as<int>(), dS.as<int>());
  disk.finish();
  std::vector<double> T((size_t)n * s.npairs), L((size_t)n * nout);
  HN_CHECK(hipMemcpy(T.data(), dTp.p, sizeof(double) * T.size(), hipMemcpyDeviceToHost));
  HN_CHECK(hipMemcpy(L.data(), dL.p, sizeof(double) * L.size(), hipMemcpyDeviceToHost));
  HN_CHECK(hipMemcpy(s.item_status.data(), dS.p, sizeof(int) * n, hipMemcpyDeviceToHost));
  if (s.ncont > 0)
    HN_CHECK(hipMemcpy(s.item_ncyc.data(), dN.p, sizeof(int) * n * s.ncont, hipMemcpyDeviceToHost));
  for (int ip = 0; ip < s.npairs; ++ip)
    for (int i = 0; i < n; ++i) s.tunn[(size_t)ip * n + i] = T[(size_t)i * s.npairs + ip] * s.p.kwght;
  s.ldos_cols = s.ndos;
  s.ldos.assign((size_t)n * s.ndos, 0.0);
  for (int l = 0; l < nout; ++l)
    for (int i = 0; i < n; ++i) s.ldos[(size_t)l * n + i] = L[(size_t)i * nout + l] * s.p.kwght;
}

// compute_landauer: extract_cont + tunneling_int_def + tunneling_and_dos
void compute_landauer(Session &s) {
  const std::vector<int> ni = pairs_of(s.p.ni), nf = pairs_of(s.p.nf);
  s.npairs = (int)ni.size();
  s.tunn.clear();
  if (s.p.emax <= s.p.emin) {  // tunneling_and_dos early return, integrations.F90:1695-1701
    s.npoints = 0;
    s.en.clear();
    s.currents.assign(1, 0.0);
    s.npairs = 1;
    return;
  }
  energy_grid(s);
  s.tunn.assign((size_t)s.npoints * s.npairs, 0.0);
  s.item_status.assign(s.npoints, 0);
  s.item_ncyc.assign((size_t)s.npoints * std::max(s.ncont, 1), 0);
  if (s.ndos > 0) {  // do_ledos: calculate_transmissions_and_dos (integrations.F90:1758-1766)
    compute_landauer_dos(s, ni, nf);
    return;
  }
  if (s.ncont < 2) return;  // calculate_transmissions: one contact -> T = 0
  hn::DevStruct *st = &device_struct(s);
  hn::Engine &eng = engine(s);
  const int n = s.npoints;
  DevBuf dE(sizeof(double) * n), dsl(sizeof(int) * n), dN(sizeof(int) * n * s.ncont),
      dS(sizeof(int) * n);
  HN_CHECK(hipMemcpy(dE.p, s.en.data(), sizeof(double) * n, hipMemcpyHostToDevice));
  HN_CHECK(hipMemset(dsl.p, 0, sizeof(int) * n));
  DevBuf dTp(sizeof(double) * n * s.npairs);
  hn::SolveOpts o = opts_of(s);
  DiskSgf disk(s, *st, eng, o);
  // every lead pair in one batched solve: contacts and gsmr blocks are shared
  eng.transmission_pairs(*st, n, dE.as<double>(), dsl.as<int>(), o, s.npairs, ni.data(), nf.data(),
                         dTp.as<double>(), dN.as<int>(), dS.as<int>());
  disk.finish();
  std::vector<double> T((size_t)n * s.npairs);
  HN_CHECK(hipMemcpy(T.data(), dTp.p, sizeof(double) * n * s.npairs, hipMemcpyDeviceToHost));
  HN_CHECK(hipMemcpy(s.item_status.data(), dS.p, sizeof(int) * n, hipMemcpyDeviceToHost));
  HN_CHECK(hipMemcpy(s.item_ncyc.data(), dN.p, sizeof(int) * n * s.ncont, hipMemcpyDeviceToHost));
  for (int ip = 0; ip < s.npairs; ++ip)
    for (int i = 0; i < n; ++i)
      s.tunn[(size_t)ip * n + i] = T[(size_t)i * s.npairs + ip] * s.p.kwght;  // tunn_mat(i,ip) = T kwght
}

// compute_dephasing_transmission: Meir-Wingreen with occupations (1, 0),
// then tunn_mat = curr_mat(:, 1) (libnegf.F90:1940-1963)
void compute_dephasing_transmission(Session &s) {
  if (s.ncont != 2) throw std::runtime_error("Effective transmission is only supported for 2 electrodes");
  const std::vector<int> ni = pairs_of(s.p.ni), nf = pairs_of(s.p.nf);
  energy_grid(s);
  const int n = s.npoints;
  std::vector<double> frm(s.ncont + 1, 0.0);
  frm[ni[0] - 1] = 1.0;
  frm[nf[0] - 1] = 0.0;
  int pad = 0;
  hn::DevStruct *st = &scba_struct(s, pad);
  const hn::DevStruct &phys = *s.dev;
  int ndev = 0;
  for (int b : phys.bsz) ndev += b;
  if ((int)s.coupling.size() < ndev) throw std::runtime_error("dephasing coupling shorter than the device");
  // the coupling in the device layout the SCBA runs on (zero on padding orbitals)
  std::vector<double> cpl(s.coupling.begin(), s.coupling.begin() + ndev);
  if (pad > 0) {
    cpl.assign((size_t)pad * phys.nbl, 0.0);
    for (int i = 0, o0 = 0; i < phys.nbl; o0 += phys.bsz[i], ++i)
      for (int r = 0; r < phys.bsz[i]; ++r) cpl[(size_t)i * pad + r] = s.coupling[o0 + r];
  }
  const int ndev_run = (int)cpl.size();
  hn::Engine &eng = engine(s);
  DevBuf dE(sizeof(double) * n), dsl(sizeof(int) * n), dC(sizeof(double) * n * s.ncont),
      dI(sizeof(int) * n), dErr(sizeof(double) * n), dS(sizeof(int) * n),
      dCp(sizeof(double) * ndev_run), dF(sizeof(double) * (s.ncont + 1));
  HN_CHECK(hipMemcpy(dE.p, s.en.data(), sizeof(double) * n, hipMemcpyHostToDevice));
  HN_CHECK(hipMemset(dsl.p, 0, sizeof(int) * n));
  HN_CHECK(hipMemcpy(dCp.p, cpl.data(), sizeof(double) * ndev_run, hipMemcpyHostToDevice));
  HN_CHECK(hipMemcpy(dF.p, frm.data(), sizeof(double) * (s.ncont + 1), hipMemcpyHostToDevice));
  hn::SCBAOpts so;
  so.coupling = dCp.as<double>();
  so.niter = s.scba_niter;
  so.tol = s.p.scba_elastic_tol;
  so.frm = dF.as<double>();
  so.frm_per_item = 0;
  so.ref = s.ncont + 1;
  hn::SolveOpts o = opts_of(s);
  DiskSgf disk(s, *st, eng, o);
  eng.scba_mw(*st, n, dE.as<double>(), dsl.as<int>(), o, so, dC.as<double>(), dI.as<int>(),
              dErr.as<double>(), dS.as<int>(), nullptr, nullptr);
  disk.finish();
  std::vector<double> cm((size_t)n * s.ncont);
  HN_CHECK(hipMemcpy(cm.data(), dC.p, sizeof(double) * n * s.ncont, hipMemcpyDeviceToHost));
  s.item_status.assign(n, 0);
  HN_CHECK(hipMemcpy(s.item_status.data(), dS.p, sizeof(int) * n, hipMemcpyDeviceToHost));
  s.curr_mat.assign((size_t)n * s.ncont, 0.0);
  for (int i = 0; i < n; ++i)
    for (int c = 0; c < s.ncont; ++c)
      s.curr_mat[(size_t)c * n + i] = cm[(size_t)i * s.ncont + c] * s.p.kwght;  // column-major
  s.npairs = 1;
  s.tunn.assign(s.curr_mat.begin(), s.curr_mat.begin() + n);
}

// ---------------------------------------------------------------------------
// Density matrix: compute_density_dft (libnegf.F90:1664-1679) = extract_cont,
// set_ref_cont (libnegf.F90:2132-2157) and contour_driver
// (integrations.F90:597-646) without inelastic terms: contour_int_def +
// contour_int (:675-897) and real_axis_int_def + real_axis_int (:989-1176).
// The per-point block work runs batched on the engine (Engine::density); the
// host builds the energy grids and assembles the CSR result.
// ---------------------------------------------------------------------------
const double kPi = 3.14159265358979323846;

// gauleg, integrations.F90:1468-1526
void gauleg(double x1, double x2, std::vector<double> &x, std::vector<double> &w, int n) {
  x.assign(n, 0.0);
  w.assign(n, 0.0);
  const int m = (n + 1) / 2;
  const double xm = 0.5 * (x2 + x1), xl = 0.5 * (x2 - x1);
  for (int i = 1; i <= m; ++i) {
    double z = std::cos(kPi * (i - 0.25) / (n + 0.5)), z1, pp, p0, p1, p2;
    for (;;) {
      p0 = 1.0;
      p1 = 0.0;
      for (int k = 1; k <= n; ++k) {
        p2 = p1;
        p1 = p0;
        p0 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p2) / k;
      }
      pp = n * (z * p0 - p1) / (z * z - 1.0);
      z1 = z;
      z = z1 - p0 / pp;
      if (std::fabs(z - z1) <= 1e-15) break;
    }
    x[i - 1] = xm - xl * z;
    x[n - i] = xm + xl * z;
    w[i - 1] = 2.0 * xl / ((1.0 - z * z) * pp * pp);
    w[n - i] = w[i - 1];
  }
}

typedef std::complex<double> cplx;

// fermi_fc, distributions.F90:77-110
cplx fermi_c(cplx E, double Ef, double kT) {
  if (kT == 0.0) return E.real() > Ef ? cplx(0.0) : cplx(1.0);
  if (std::fabs((E.real() - Ef) / kT) > 30)
    return E.real() > Ef ? std::exp(-(E - Ef) / kT) : cplx(1.0);
  return 1.0 / (std::exp((E - Ef) / kT) + 1.0);
}

// contour_int_def without inelastic terms (integrations.F90:675-835)
void contour_grid(double Elow, double muref, double kbT, int n_kt, int np1, int np2, int npoles,
                  double g_spin, std::vector<cplx> &z, std::vector<cplx> &w) {
  const cplx I(0.0, 1.0);
  const double nkT = n_kt * kbT, mumin = muref - nkT;
  double Lambda = 2.0 * npoles * kbT * kPi;
  if (npoles == 0) Lambda = 0.5 * kbT * kPi;
  const double Centre = (Lambda * Lambda - Elow * Elow + mumin * mumin) / (2.0 * (mumin - Elow));
  const double Rad = Centre - Elow;
  const double alpha = (kbT == 0.0) ? 0.1 * kPi : std::atan(Lambda / (mumin - Centre));
  std::vector<double> x, wt;
  z.clear();
  w.clear();
  gauleg(kPi, alpha, x, wt, np1);
  for (int i = 0; i < np1; ++i) {
    const cplx Pc = Rad * std::exp(I * x[i]);
    z.push_back(Centre + Pc);
    w.push_back(I * Pc * g_spin * wt[i] / (2.0 * kPi));
  }
  cplx z1 = 0.0, zdiff = 0.0;
  if (kbT == 0.0) {
    gauleg(alpha, 0.0, x, wt, np2);
  } else {
    z1 = cplx(muref + nkT, Lambda);
    const cplx z2(muref - nkT, Lambda);
    zdiff = z2 - z1;
    gauleg(1.0, 0.0, x, wt, np2);
  }
  for (int i = 0; i < np2; ++i) {
    if (kbT == 0.0) {
      const cplx Pc = Rad * std::exp(I * x[i]);
      z.push_back(Centre + Pc);
      const double dt = g_spin * wt[i] / (2.0 * kPi);
      w.push_back(dt * Pc * I);
    } else {
      const cplx Ec = z1 + x[i] * zdiff;
      z.push_back(Ec);
      w.push_back(g_spin * zdiff * fermi_c(Ec, muref, kbT) * wt[i] / (2.0 * kPi));
    }
  }
  for (int i = 1; i <= npoles; ++i) {
    z.push_back(cplx(muref, kbT * kPi * (2.0 * i - 1.0)));
    w.push_back(-I * g_spin * kbT);
  }
}

// Sparse accumulation of (row, col, value) entries; duplicates are summed.
struct SparseAcc {
  std::vector<std::vector<std::pair<int, double2>>> rows;
  explicit SparseAcc(int n) : rows(n) {}
  void add(int r, int c, double2 v) { rows[r].push_back({c, v}); }
  void finalize() {
    for (auto &row : rows) {
      std::stable_sort(row.begin(), row.end(), [](const std::pair<int, double2> &a,
                                           const std::pair<int, double2> &b) { return a.first < b.first; });
      size_t o = 0;
      for (size_t i = 0; i < row.size(); ++i) {
        if (o > 0 && row[o - 1].first == row[i].first) {
          row[o - 1].second.x += row[i].second.x;
          row[o - 1].second.y += row[i].second.y;
        } else {
          row[o++] = row[i];
        }
      }
      row.resize(o);
    }
  }
  double2 at(int r, int c) const {
    const auto &row = rows[r];
    auto it = std::lower_bound(row.begin(), row.end(), c, [](const std::pair<int, double2> &a, int v) {
      return a.first < v;
    });
    return (it != row.end() && it->first == c) ? it->second : make_double2(0.0, 0.0);
  }
};

// One integration path's accumulated blocks put on the density-matrix pattern:
// the full diagonal (the zero-initialised CSR of contour_int / real_axis_int),
// blk2csr of the device blocks on the S pattern (iterative.F90:1375-1450:
// values on device rows and columns, explicit zeros elsewhere), and the outer
// blocks on the coupling pattern (HMC or SMC nonzero) at (block cb rows,
// contact PL1 columns) and, for outer == 2, at the transposed position.
SparseAcc place_blocks(Session &s, const hn::DevStruct &st, const std::vector<double2> &acc,
                       const std::vector<long long> &lay, int N, bool lower) {
  SparseAcc A(N);
  const int nbl = st.nbl, ncont = st.ncont, ndev = st.ndev();
  std::vector<int> start(nbl + 1, 0), blk_of(ndev);
  for (int i = 0; i < nbl; ++i) start[i + 1] = start[i] + st.bsz[i];
  for (int i = 0; i < nbl; ++i)
    for (int r = start[i]; r < start[i + 1]; ++r) blk_of[r] = i;
  const double2 z0 = make_double2(0.0, 0.0);
  for (int r = 0; r < N; ++r) A.add(r, r, z0);
  auto blockval = [&](int r, int c) -> double2 {
    const int x = blk_of[r], y = blk_of[c];
    const int rr = r - start[x], cc = c - start[y];
    if (y == x) return acc[lay[x] + rr + (long long)cc * st.bsz[x]];
    if (y == x + 1) return acc[lay[nbl + x] + rr + (long long)cc * st.bsz[x]];
    if (y == x - 1) return acc[lay[2 * nbl - 1 + y] + rr + (long long)cc * st.bsz[x]];
    throw std::runtime_error("blk2csr: overlap pattern outside the block tridiagonal (wrong PL size?)");
  };
  const bool sid = s.cur >= (int)s.S.size() || s.s_id[s.cur] || !s.S[s.cur].present();
  if (sid) {
    for (int r = 0; r < N; ++r) A.add(r, r, r < ndev ? blockval(r, r) : z0);
  } else {
    const Csr &S = s.S[s.cur];
    if (S.n != N) throw std::runtime_error("S and H sizes differ");
    for (int r = 0; r < N; ++r)
      for (int k = S.ptr[r]; k < S.ptr[r + 1]; ++k) {
        const int c = S.col[k];
        A.add(r, c, (r < ndev && c < ndev) ? blockval(r, c) : z0);
      }
  }
  const Csr &H = s.H[s.cur];
  for (int k = 0; k < ncont; ++k) {
    const int cb = st.cblk[k] - 1, nd = st.bsz[cb], nc = st.ngs(k);  // surface + PL1 columns
    const int r0 = start[cb], c0 = s.surfstart[k] - 1;
    const std::vector<double2> hm = extract(H, r0 + 1, r0 + nd, c0 + 1, c0 + nc);
    std::vector<double2> sm;
    if (!sid) sm = extract(s.S[s.cur], r0 + 1, r0 + nd, c0 + 1, c0 + nc);
    const long long olc = lay[3 * nbl - 2 + k], ocl = lay[3 * nbl - 2 + ncont + k];
    for (int j = 0; j < nc; ++j)
      for (int i = 0; i < nd; ++i) {
        const size_t t = i + (size_t)j * nd;
        const bool nz = hm[t].x != 0.0 || hm[t].y != 0.0 || (!sid && (sm[t].x != 0.0 || sm[t].y != 0.0));
        if (!nz) continue;
        A.add(r0 + i, c0 + j, acc[olc + t]);
        if (lower) A.add(c0 + j, r0 + i, acc[ocl + j + (size_t)i * nc]);
      }
  }
  A.finalize();
  return A;
}

// set_ref_cont (libnegf.F90:2132-2157): reference contact and its mu
void set_ref_cont(const Session &s, int &refcont, double &muref) {
  int imin = 0, imax = 0;
  for (int c = 1; c < s.ncont; ++c) {
    if (s.p.mu[c] < s.p.mu[imin]) imin = c;
    if (s.p.mu[c] > s.p.mu[imax]) imax = c;
  }
  if (s.p.min_or_max == 0) { muref = s.p.mu[imin]; refcont = imin + 1; }
  else if (s.p.min_or_max == 1) { muref = s.p.mu[imax]; refcont = imax + 1; }
  else { muref = s.p.mu[imax]; refcont = s.ncont + 1; }
}

// solve options of the density paths: readOldDM_SGFs selects the cache use
hn::SolveOpts dm_opts(Session &s) {
  hn::SolveOpts o = opts_of(s);
  const int r = s.p.readold_dm_sgfs;
  o.sgf_mode = s.p.sgfcache == 2 ? 0 : (r == 0) ? 1 : (r == 2) ? 2 : 0;
  if (o.sgf_mode && !s.sgf_cache_on) {
    engine(s).set_sgf_cache(2ull << 30);
    s.sgf_cache_on = true;
  }
  return o;
}

// rho += i (T - T^H) (zspectral) or rho += sign T, over T's entries
void add_spectral(SparseAcc &rho, const SparseAcc &T, int N) {
  for (int r = 0; r < N; ++r)
    for (const auto &e : T.rows[r]) {
      const double2 v = e.second;
      rho.add(r, e.first, make_double2(-v.y, v.x));   // i T(r,c)
      rho.add(e.first, r, make_double2(-v.y, -v.x));  // -i conj(T(r,c)) at (c,r)
    }
}
void add_plain(SparseAcc &rho, const SparseAcc &T, int N, double sign) {
  for (int r = 0; r < N; ++r)
    for (const auto &e : T.rows[r]) rho.add(r, e.first, make_double2(sign * e.second.x, sign * e.second.y));
}

// negf%rho as 0-based CSR
void store_rho(Session &s, SparseAcc &rho, int N) {
  rho.finalize();
  s.rho_n = N;
  s.rho_ptr.assign(N + 1, 0);
  s.rho_col.clear();
  s.rho_val.clear();
  for (int r = 0; r < N; ++r) {
    for (const auto &e : rho.rows[r]) {
      s.rho_col.push_back(e.first);
      s.rho_val.push_back(e.second);
    }
    s.rho_ptr[r + 1] = (int)s.rho_col.size();
  }
}

// One batched pass of the density engine over the points z with weights w
// (mode 0: sum w Gr on the pattern; 1: sum w Gn with occupations frm).
std::vector<double2> density_run(Session &s, hn::DevStruct &st, hn::Engine &eng, const hn::SolveOpts &o,
                               long long tot, int refcont, int mode, const std::vector<cplx> &z,
                               const std::vector<cplx> &w, const std::vector<double> &frm,
                               bool real_device = false) {
  const int n = (int)z.size();
  std::vector<double> Er(n), Ei(n);
  std::vector<double2> ww(n);
  for (int i = 0; i < n; ++i) {
    Er[i] = z[i].real();
    Ei[i] = z[i].imag();
    ww[i] = make_double2(w[i].real(), w[i].imag());
  }
  DevBuf dE(sizeof(double) * n), dEi(sizeof(double) * n), dsl(sizeof(int) * n), dw(sizeof(double2) * n),
      dF(sizeof(double) * std::max<size_t>(frm.size(), 1)), dA(sizeof(double2) * std::max<long long>(tot, 1)),
      dS(sizeof(int) * n);
  HN_CHECK(hipMemcpy(dE.p, Er.data(), sizeof(double) * n, hipMemcpyHostToDevice));
  HN_CHECK(hipMemcpy(dEi.p, Ei.data(), sizeof(double) * n, hipMemcpyHostToDevice));
  HN_CHECK(hipMemset(dsl.p, 0, sizeof(int) * n));
  HN_CHECK(hipMemcpy(dw.p, ww.data(), sizeof(double2) * n, hipMemcpyHostToDevice));
  if (!frm.empty()) HN_CHECK(hipMemcpy(dF.p, frm.data(), sizeof(double) * frm.size(), hipMemcpyHostToDevice));
  HN_CHECK(hipMemset(dA.p, 0, sizeof(double2) * tot));
  hn::DensityOpts d;
  d.mode = mode;
  d.w = dw.as<double2>();
  d.frm = mode == 1 ? dF.as<double>() : nullptr;
  d.ref = refcont;
  d.outer = s.outer;
  d.status = dS.as<int>();
  d.real_device = real_device ? 1 : 0;
  eng.density(st, n, dE.as<double>(), mode == 0 ? dEi.as<double>() : nullptr, dsl.as<int>(), o, d,
              dA.as<double2>());
  std::vector<double2> acc(tot);
  HN_CHECK(hipMemcpy(acc.data(), dA.p, sizeof(double2) * tot, hipMemcpyDeviceToHost));
  std::vector<int> stv(n);
  HN_CHECK(hipMemcpy(stv.data(), dS.p, sizeof(int) * n, hipMemcpyDeviceToHost));
  s.item_status.insert(s.item_status.end(), stv.begin(), stv.end());
  return acc;
}

// compute_ldos (libnegf.F90:1893-1901) = extract_cont + tunneling_int_def +
// ldos_int (integrations.F90:212-258): per point of the transmission grid,
// G^r at Ec = E + i dos_delta (device and contacts), with the device/contact
// blocks (compute_Gr(negf, outer = 1), G^r kept on the pattern of S), then
// ldos_mat(i, l) = sum over the orbitals of projection l of
// -g_spin Im(diag(G^r S)) / pi.  Orbitals past the device have no G^r rows
// (their diagonal is 0).  The elastic SCBA loop of compute_Gr is not run
// here: with elastic interactions and scba_niter > 0 this fails.
void compute_ldos(Session &s) {
  if (s.deph && s.scba_niter > 0)
    throw std::runtime_error("compute_ldos: elastic interactions (scba_niter > 0) are not supported");
  if (s.ndos < 1) throw std::runtime_error("compute_ldos: no LDOS projections (init_ldos)");
  energy_grid(s);
  const int n = s.npoints;
  s.ldos_cols = s.ndos;
  s.ldos.assign((size_t)n * s.ndos, 0.0);
  s.item_status.assign(n, 0);
  if (n <= 0) return;
  hn::DevStruct *st = &device_struct(s);
  hn::Engine &eng = engine(s);
  const int ndev = st->ndev(), N = s.H[s.cur].n;
  for (int l = 0; l < s.ndos; ++l)
    for (int i2 : s.dos_proj[l])
      if (i2 < 1 || i2 > N) throw std::runtime_error("LDOS orbital index outside 1..size of H");
  const hn::SolveOpts o = dm_opts(s);
  std::vector<double> Ei(n, s.p.dos_delta);
  DevBuf dE(sizeof(double) * n), dEi(sizeof(double) * n), dsl(sizeof(int) * n), dS(sizeof(int) * n),
      dG(sizeof(double) * (size_t)n * ndev);
  HN_CHECK(hipMemcpy(dE.p, s.en.data(), sizeof(double) * n, hipMemcpyHostToDevice));
  HN_CHECK(hipMemcpy(dEi.p, Ei.data(), sizeof(double) * n, hipMemcpyHostToDevice));
  HN_CHECK(hipMemset(dsl.p, 0, sizeof(int) * n));
  HN_CHECK(hipMemset(dG.p, 0, sizeof(double) * (size_t)n * ndev));
  hn::DensityOpts d;
  d.mode = 0;
  d.outer = 1;
  d.status = dS.as<int>();
  d.gs_diag = dG.as<double>();
  d.gs_ld = ndev;
  d.gs_scale = -s.p.g_spin / kPi;
  eng.density(*st, n, dE.as<double>(), dEi.as<double>(), dsl.as<int>(), o, d, nullptr);
  std::vector<double> g((size_t)n * ndev);
  HN_CHECK(hipMemcpy(g.data(), dG.p, sizeof(double) * g.size(), hipMemcpyDeviceToHost));
  HN_CHECK(hipMemcpy(s.item_status.data(), dS.p, sizeof(int) * n, hipMemcpyDeviceToHost));
  for (int l = 0; l < s.ndos; ++l)
    for (int i = 0; i < n; ++i) {
      double v = 0.0;
      for (int i2 : s.dos_proj[l])
        if (i2 <= ndev) v += g[(size_t)i * ndev + i2 - 1];
      s.ldos[(size_t)l * n + i] = v;
    }
}

void compute_density(Session &s) {
  if (s.ncont < 1) throw std::runtime_error("density matrix needs at least one contact");
  hn::DevStruct *st = &device_struct(s);
  hn::Engine &eng = engine(s);
  const int ncont = s.ncont, N = s.H[s.cur].n;
  int refcont;
  double muref;
  set_ref_cont(s, refcont, muref);
  const bool energy = s.p.dore[0] == 'E';
  hn::SolveOpts o = dm_opts(s);
  const std::vector<long long> lay = hn::Engine::density_layout(*st);
  const long long tot = lay.back();
  SparseAcc rho(N);
  s.item_status.clear();
  auto run = [&](int mode, const std::vector<cplx> &z, const std::vector<cplx> &w,
                 const std::vector<double> &frm) {
    return density_run(s, *st, eng, o, tot, refcont, mode, z, w, frm);
  };
  // en_grid as the reference leaves it: the last integration path's points
  s.en.clear();
  s.en_im.clear();
  // ---- contour_int: rho = i (T - T^H), T = sum w Gr on the pattern ----
  if (s.p.np_n[0] + s.p.np_n[1] + s.p.n_poles > 0) {
    // kbT of the reference contact (the contact of max mu when there is none)
    const int kc = refcont <= ncont ? refcont - 1 : [&] {
      int im = 0;
      for (int c = 1; c < ncont; ++c) if (s.p.mu[c] > s.p.mu[im]) im = c;
      return im;
    }();
    std::vector<cplx> z, w;
    contour_grid(s.p.ec, muref, s.p.kbt_dm[kc], s.p.n_kt, s.p.np_n[0], s.p.np_n[1], s.p.n_poles,
                 s.p.g_spin, z, w);
    if (energy)
      for (size_t i = 0; i < z.size(); ++i) w[i] *= z[i];
    for (const cplx &v : z) {
      s.en.push_back(v.real());
      s.en_im.push_back(v.imag());
    }
    const std::vector<double2> acc = run(0, z, w, {});
    add_spectral(rho, place_blocks(s, *st, acc, lay, N, s.outer == 2), N);
  }
  // ---- real_axis_int: rho += sum w Gn on the pattern ----
  if (s.p.np_real > 0) {
    double mumin = s.p.mu[0], mumax = s.p.mu[0], kmax = s.p.kbt_dm[0];
    for (int c = 1; c < ncont; ++c) {
      mumin = std::min(mumin, s.p.mu[c]);
      mumax = std::max(mumax, s.p.mu[c]);
      kmax = std::max(kmax, s.p.kbt_dm[c]);
    }
    const double Omega = s.p.n_kt * kmax;
    std::vector<double> x, wt;
    gauleg(mumin - Omega, mumax + Omega, x, wt, s.p.np_real);
    const int n = s.p.np_real;
    std::vector<cplx> z(n), w(n);
    std::vector<double> frm((size_t)n * (ncont + 1), 0.0);
    s.en.clear();
    s.en_im.clear();
    for (int i = 0; i < n; ++i) {
      z[i] = cplx(x[i], s.p.delta);
      w[i] = s.p.g_spin * wt[i] / (2.0 * kPi);
      if (energy) w[i] *= x[i];
      for (int c = 0; c < ncont; ++c) frm[(size_t)i * (ncont + 1) + c] = fermi(x[i], s.p.mu[c], s.p.kbt_dm[c]);
      s.en.push_back(x[i]);
      s.en_im.push_back(s.p.delta);
    }
    o.delta = s.p.delta;
    const std::vector<double2> acc = run(1, z, w, frm);
    add_plain(rho, place_blocks(s, *st, acc, lay, N, s.outer == 2), N, 1.0);
  }
  store_rho(s, rho, N);
}

// trapez, integrations.F90:1529-1547
void trapez(double x1, double x2, std::vector<double> &x, std::vector<double> &w, int n) {
  if (n < 2) throw std::runtime_error("trapezoid rule needs at least 2 points (np_real)");
  const double d = (x2 - x1) / (double)(n - 1);
  x.assign(n, 0.0);
  w.assign(n, d);
  w[0] = w[n - 1] = 0.5 * d;
  for (int i = 1; i <= n; ++i) x[i - 1] = (x1 * (n - i) + x2 * (i - 1)) / (n - 1);
}

// Electron (particle 1) or hole (-1) contour of contour_int_n_def /
// contour_int_p_def (integrations.F90:281-426, 447-592): from the band edge
// Eb (Ec - DeltaEc, or Ev + DeltaEv) up to Eb + i Lambda, across to
// muref +- Omega + i Lambda, then the Fermi poles (none when the edge lies
// beyond muref); holes weight 1 - f and run the other way.
void efa_contour_grid(int particle, double Eb, double muref, double kbT, int n_kt, int np1, int np2,
                      int npoles, double g_spin, std::vector<cplx> &z, std::vector<cplx> &w) {
  const cplx I(0.0, 1.0);
  const double Omega = n_kt * kbT;
  const double Lambda = npoles == 0 ? 0.5 * kbT * kPi : 2.0 * npoles * kbT * kPi;
  if (Eb < muref + 0.001 && Eb > muref - 0.001) Eb = muref - particle * kbT;
  const int Np = (particle == 1 ? Eb > muref : Eb < muref) ? 0 : npoles;
  z.clear();
  w.clear();
  cplx z1 = Eb, z2 = Eb + I * Lambda;
  for (int seg = 0; seg < 2; ++seg) {
    if (seg == 1) {
      z1 = z2;
      z2 = muref + particle * Omega + I * Lambda;
    }
    const cplx zd = z2 - z1;
    std::vector<double> x, wt;
    gauleg(0.0, 1.0, x, wt, seg == 0 ? np1 : np2);
    for (size_t i = 0; i < x.size(); ++i) {
      const cplx Ec = z1 + x[i] * zd;
      const cplx ff = particle == 1 ? fermi_c(Ec, muref, kbT) : fermi_c(-Ec, -muref, kbT);
      z.push_back(Ec);
      w.push_back((double)particle * g_spin * zd * ff * wt[i] / (2.0 * kPi));
    }
  }
  for (int i = 1; i <= Np; ++i) {
    z.push_back(muref + I * kbT * kPi * (2.0 * i - 1.0));
    w.push_back(-(double)particle * I * kbT * g_spin);
  }
}

// compute_density_efa (libnegf.F90:1690-1791) without the bulk contact
// density: the electron (hole) contour with muref = mu_n (mu_p) of contact 1,
// then the real-axis trapezoid of real_axis_int_{n,p}_def
// (integrations.F90:1346-1462) through real_axis_int (:1093-1176), whose
// hole integrand is compute_Gp = A - Gn with A = i (Gr - Gr^H) (:2405-2420);
// q(k) = Re rho(k, k).  rho is kept for negf_get_dm.
void compute_density_efa(Session &s, int particle, double *q, int ndofs) {
  if (particle != 1 && particle != -1)
    throw std::runtime_error("libNEGF error. In compute_density_efa, unknown particle");
  if (s.ncont < 1) throw std::runtime_error("density needs at least one contact");
  hn::DevStruct *st = &device_struct(s);
  hn::Engine &eng = engine(s);
  const int ncont = s.ncont, N = s.H[s.cur].n;
  if (ndofs > N) throw std::runtime_error("compute_density_efa: ndofs larger than H");
  int refcont;
  double muref;
  set_ref_cont(s, refcont, muref);
  muref = particle == 1 ? s.p.mu_n[0] : s.p.mu_p[0];
  const int *np = particle == 1 ? s.p.np_n : s.p.np_p;
  const bool contour = np[0] + np[1] + s.p.n_poles > 0;
  if (!contour) refcont = ncont + 1;
  const bool energy = s.p.dore[0] == 'E';
  hn::SolveOpts o = dm_opts(s);
  const std::vector<long long> lay = hn::Engine::density_layout(*st);
  const long long tot = lay.back();
  auto run = [&](int mode, const std::vector<cplx> &z, const std::vector<cplx> &w,
                 const std::vector<double> &frm, bool real_device = false) {
    return place_blocks(s, *st, density_run(s, *st, eng, o, tot, refcont, mode, z, w, frm, real_device),
                        lay, N, s.outer == 2);
  };
  SparseAcc rho(N);
  s.item_status.clear();
  s.en.clear();
  s.en_im.clear();
  bool any = false;
  if (contour) {
    // kbT of the reference contact (the contact of max mu when there is none)
    int kc = refcont - 1;
    if (refcont > ncont) {
      kc = 0;
      for (int c = 1; c < ncont; ++c) if (s.p.mu[c] > s.p.mu[kc]) kc = c;
    }
    std::vector<cplx> z, w;
    const double Eb = particle == 1 ? s.p.ec - s.p.deltaec : s.p.ev + s.p.deltaev;
    efa_contour_grid(particle, Eb, muref, s.p.kbt_dm[kc], s.p.n_kt, np[0], np[1], s.p.n_poles, s.p.g_spin,
                     z, w);
    if (energy)
      for (size_t i = 0; i < z.size(); ++i) w[i] *= z[i];
    for (const cplx &v : z) {
      s.en.push_back(v.real());
      s.en_im.push_back(v.imag());
    }
    if (!z.empty()) {
      add_spectral(rho, run(0, z, w, {}), N);
      any = true;
    }
  }
  if (s.p.np_real > 0) {
    double kmax = s.p.kbt_dm[0], mun = s.p.mu_n[0], mup = s.p.mu_p[0];
    for (int c = 1; c < ncont; ++c) {
      kmax = std::max(kmax, s.p.kbt_dm[c]);
      mun = std::max(mun, s.p.mu_n[c]);
      mup = std::min(mup, s.p.mu_p[c]);
    }
    const double Omega = s.p.n_kt * kmax;
    std::vector<double> x, wt;
    if (particle == 1) trapez(s.p.ec - s.p.deltaec, mun + Omega, x, wt, s.p.np_real);
    else trapez(mup - Omega, s.p.ev + s.p.deltaev, x, wt, s.p.np_real);
    const int n = s.p.np_real;
    std::vector<cplx> z(n), w(n);
    std::vector<double> frm((size_t)n * (ncont + 1), 0.0);
    s.en.clear();
    s.en_im.clear();
    for (int i = 0; i < n; ++i) {
      z[i] = cplx(x[i], s.p.delta);
      w[i] = s.p.g_spin * wt[i] / (2.0 * kPi);
      if (energy) w[i] *= x[i];
      for (int c = 0; c < ncont; ++c) frm[(size_t)i * (ncont + 1) + c] = fermi(x[i], s.p.mu[c], s.p.kbt_dm[c]);
      s.en.push_back(x[i]);
      s.en_im.push_back(s.p.delta);
    }
    o.delta = s.p.delta;
    if (particle == 1) {
      add_plain(rho, run(1, z, w, frm), N, 1.0);
    } else {  // Gp = A - Gn, A from the G^r that comes with G^n (device at real E)
      add_spectral(rho, run(0, z, w, {}, true), N);
      add_plain(rho, run(1, z, w, frm), N, -1.0);
    }
    any = true;
  }
  store_rho(s, rho, N);
  for (int k = 0; k < ndofs; ++k) {
    double v = 0.0;
    if (any)
      for (int i = s.rho_ptr[k]; i < s.rho_ptr[k + 1]; ++i)
        if (s.rho_col[i] == k) v += s.rho_val[i].x;
    q[k] = v;
  }
}

// compute_current, libnegf.F90:1904-1919
void compute_current(Session &s) {
  if (s.deph) compute_dephasing_transmission(s);
  else compute_landauer(s);
  if (!s.tunn.empty()) electron_current(s);
}

// convertCurrent, ln_constants.F90:167-194 (tables :88-106)
double convert_current(const std::string &h, const std::string &j) {
  const double HAR = 27.21138506, eovh = 1.054181532603995e-3, Kb = 3.166811389002312e-6;
  const double J_eV = 1.0 / 1.60217656e-19;
  if (h == "unknown") return 1.0;
  double curr = std::nan("");
  struct U { const char *n; double v; };
  const U eu[] = {{"unknown", 1.0}, {"au", 1.0}, {"H", 1.0}, {"Ry", 0.5}, {"eV", 1.0 / HAR},
                  {"kcal/mol", 0.0433634 / HAR}, {"K", Kb}, {"cm^-1", 1.239841930e-4 / HAR},
                  {"J", J_eV / HAR}};
  const U cu[] = {{"unknown", 1.0}, {"A", 1.0}, {"mA", 1.0e-3}, {"nA", 1.0e-9}};
  for (const U &u : eu)
    if (h == u.n) curr = eovh * u.v;
  for (const U &u : cu)
    if (j == u.n) curr = curr * u.v;
  return curr;
}

std::string cstr(const char *c) {
  std::string s = c ? c : "";
  while (!s.empty() && (s.back() == ' ' || s.back() == '\0')) s.pop_back();
  return s;
}

}  // namespace

extern "C" {

void negf_gethandlersize(int *handlerSize) { *handlerSize = NEGF_HSIZE; }

void negf_init_session(int *handler) {
  static_assert(sizeof(Session *) <= NEGF_HSIZE * sizeof(int), "handler too small");
  std::memset(handler, 0, NEGF_HSIZE * sizeof(int));
  Session *s = new Session;
  set_defaults(s->p);
  std::memcpy(handler, &s, sizeof(s));
}

void negf_init(int *handler) {
  api(handler, [](Session &s) {
    set_defaults(s.p);
    s.H.assign(1, Csr());
    ++s.gen;
    s.S.assign(1, Csr());
    s.s_id.assign(1, 0);
    s.cur = 0;
  });
}

void negf_destruct_libnegf(int *handler) {
  api(handler, [](Session &s) {
    s.H.clear();
    s.S.clear();
    s.tunn.clear();
    s.curr_mat.clear();
    s.currents.clear();
    s.ldos.clear();
    s.deph = false;
    s.dev.reset();
    s.eng.reset();
    s.sgf_cache_on = false;
    ++s.gen;
  });
}

void negf_destruct_session(int *handler) {
  Session *s = sess(handler);
  delete s;
  std::memset(handler, 0, NEGF_HSIZE * sizeof(int));
}

void negf_set_verbosity(int *handler, int v) {
  api(handler, [&](Session &s) { s.p.verbose = v; });
}

void negf_init_contacts(int *handler, int ncont) {
  api(handler, [&](Session &s) {
    if (ncont < 0 || ncont > MAXNCONT) throw std::runtime_error("bad number of contacts");
    s.ncont = ncont;
    for (int c = 0; c < MAXNCONT; ++c) {
      s.p.mu[c] = s.p.mu_n[c] = s.p.mu_p[c] = 0.0;
      s.p.kbt_dm[c] = s.p.kbt_t[c] = 0.0;
      s.p.contact_dos[c] = 0.0;
      s.p.fictcont[c] = false;
    }
  });
}

void negf_init_structure(int *handler, int ncont, int *surfstart, int *surfend, int *contend,
                         int npl, int *plend, int *cblk) {
  api(handler, [&](Session &s) {
    if (s.ncont != 0 && s.ncont != ncont)
      throw std::runtime_error("Error in set_structure: ncont not compatible with previous initialization.");
    if (npl <= 0) throw std::runtime_error("automatic PL partitioning (npl = 0) is not supported");
    s.ncont = ncont;
    s.npl = npl;
    s.surfstart.assign(surfstart, surfstart + ncont);
    s.surfend.assign(surfend, surfend + ncont);
    s.contend.assign(contend, contend + ncont);
    s.cblk.assign(cblk, cblk + ncont);
    s.plend.assign(plend, plend + npl);
    s.structure_set = true;
    ++s.gen;
  });
}

// negf_contact_blocks -> find_cblocks (libnegf.F90:2172-2234): for each
// contact the device PL holding every device column that the rows
// surfstart..contend of H touch; an error when they span two PLs.
void negf_contact_blocks(int *handler, int ncont, int *surfstart, int *surfend, int *contend,
                         int npl, int *plend, int *cblk) {
  (void)surfend;
  api(handler, [&](Session &s) {
    if (s.cur >= (int)s.H.size() || !s.H[s.cur].present())
      throw std::runtime_error("Error: H not created before invoking negf_contact_block");
    if (npl <= 0) throw std::runtime_error("negf_contact_blocks: npl must be positive");
    const Csr &H = s.H[s.cur];
    const int last = plend[npl - 1];
    std::vector<int> out(ncont);
    for (int j = 0; j < ncont; ++j) {
      int mx = 0, mn = 400000000;
      for (int k = surfstart[j]; k <= contend[j]; ++k) {
        if (k < 1 || k > H.n) throw std::runtime_error("negf_contact_blocks: contact rows outside H");
        for (int i = H.ptr[k - 1]; i < H.ptr[k]; ++i) {
          const int col = H.col[i] + 1;
          if (col <= last) { mn = std::min(mn, col); mx = std::max(mx, col); }
        }
      }
      for (int k = 0; k < npl; ++k)
        if (mx <= plend[k]) {
          out[j] = k + 1;
          if (mn >= (k ? plend[k - 1] + 1 : 1)) break;
          throw std::runtime_error("ERROR: contact interacting with more than one block (contact " +
                                   std::to_string(j + 1) + ")");
        }
    }
    std::copy(out.begin(), out.end(), cblk);
  });
}

// negf_get_pls (libnegf_api.f90:245-269): npl, and with copy != 0 plend,
// ncont and cblk of the structure
void negf_get_pls(int *handler, int *npl, int *ncont, int *plend, int *cblk, int copy) {
  api(handler, [&](Session &s) {
    *npl = s.npl;
    if (!copy) return;
    std::copy(s.plend.begin(), s.plend.end(), plend);
    *ncont = s.ncont;
    std::copy(s.cblk.begin(), s.cblk.end(), cblk);
  });
}

// negf_set_reference (libnegf_api.f90:1172-1185): min_or_max of the density
// paths (0 the contact of lowest mu is the reference, 1 the highest)
void negf_set_reference(int *handler, int *minmax) {
  api(handler, [&](Session &s) { s.p.min_or_max = *minmax; });
}

// negf_print_mat (libnegf_api.f90:545-567): H.dat and S.dat in the scratch
// path as printcsr writes them (zprint_csrcoo 'c', sparsekit_drv.F90:1127-1159):
// one line per stored entry, row by row, format (2i8,2f20.10), 1-based
void negf_print_mat(int *handler) {
  api(handler, [&](Session &s) {
    if (s.cur >= (int)s.H.size() || !s.H[s.cur].present()) throw std::runtime_error("negf_print_mat: H not set");
    auto dump = [&](const std::string &fn, const Csr *A, int n) {
      FILE *f = std::fopen(fn.c_str(), "w");
      if (!f) throw std::runtime_error("cannot open " + fn);
      for (int r = 0; r < n; ++r) {
        if (A) {
          for (int i = A->ptr[r]; i < A->ptr[r + 1]; ++i)
            std::fprintf(f, "%8d%8d%20.10f%20.10f\n", r + 1, A->col[i] + 1, A->v[i].x, A->v[i].y);
        } else {
          std::fprintf(f, "%8d%8d%20.10f%20.10f\n", r + 1, r + 1, 1.0, 0.0);  // set_s_id's identity
        }
      }
      std::fclose(f);
    };
    const Csr &H = s.H[s.cur];
    dump(s.scratch + "H.dat", &H, H.n);
    const bool sid = s.cur >= (int)s.S.size() || s.s_id[s.cur] || !s.S[s.cur].present();
    dump(s.scratch + "S.dat", sid ? nullptr : &s.S[s.cur], H.n);
  });
}

// negf_write_partition -> negf_partition_info (libnegf.F90:1306-1313)
void negf_write_partition(int *handler) {
  api(handler, [&](Session &s) {
    std::printf(" (LibNEGF) Partitioning:\n Number of blocks: %12d\n Contact interactions:", s.npl);
    for (int c : s.cblk) std::printf("%12d", c);
    std::printf("\n");
    std::fflush(stdout);
  });
}

// negf_print_tnegf -> print_all_vars (lib_param.F90:674-740): the session's
// parameters, one "name= value" line each
void negf_print_tnegf(int *handler) {
  api(handler, [&](Session &s) {
    const lnparams &p = s.p;
    auto arr = [&](const char *name, const double *v) {
      std::printf(" %s=", name);
      for (int c = 0; c < std::max(s.ncont, 1); ++c) std::printf(" %.15g", v[c]);
      std::printf("\n");
    };
    std::printf(" verbose= %d\n scratch= \"%s\"\n output= \"%s\"\n", p.verbose, s.scratch.c_str(),
                s.out_path.c_str());
    arr("mu", p.mu);
    arr("mu_n", p.mu_n);
    arr("mu_p", p.mu_p);
    for (int c = 0; c < s.ncont; ++c)
      std::printf(" Contact Parameters for contact %d\n mu= %.15g\n WideBand= %s\n DOS= %.15g\n mu_n= %.15g\n"
                  " mu_p= %.15g\n Density Matrix T= %.15g\n Transmission T= %.15g\n",
                  c + 1, p.mu[c], p.fictcont[c] ? "T" : "F", p.contact_dos[c], p.mu_n[c], p.mu_p[c], p.kbt_dm[c],
                  p.kbt_t[c]);
    std::printf(" Contour Parameters:\n Ec= %.15g\n Ev= %.15g\n DEc= %.15g\n DEv= %.15g\n ReadoldDM_SGFs= %d\n"
                " ReadoldT_SGF= %d\n Np_n= %d %d\n Np_p= %d %d\n nkT= %d\n nPoles= %d\n min_or_max= %d\n",
                p.ec, p.ev, p.deltaec, p.deltaev, p.readold_dm_sgfs, p.readold_t_sgfs, p.np_n[0], p.np_n[1],
                p.np_p[0], p.np_p[1], p.n_kt, p.n_poles, p.min_or_max);
    std::printf(" Transmission Parameters:\n Emin= %.15g\n Emax= %.15g\n Estep= %.15g\n g_spin= %.15g\n"
                " delta= %.15g\n dos_delta= %.15g\n DorE= %c\n",
                p.emin, p.emax, p.estep, p.g_spin, p.delta, p.dos_delta, p.dore[0]);
    std::printf(" Internal variables:\n kp= %d\n wght= %.15g\n outer_blocks= %d\n", p.kpoint, p.kwght, s.outer);
    std::fflush(stdout);
  });
}

// negf_mem_stats (libnegf_api.f90:1462-1475; writePeakInfo / writeMemInfo):
// the device memory in use and free, and the host's peak resident size
void negf_mem_stats(int *handler) {
  api(handler, [&](Session &) {
    size_t fr = 0, tot = 0;
    HN_CHECK(hipMemGetInfo(&fr, &tot));
    long peak_kb = 0;
    if (FILE *f = std::fopen("/proc/self/status", "r")) {
      char line[256];
      while (std::fgets(line, sizeof(line), f))
        if (std::strncmp(line, "VmHWM:", 6) == 0) peak_kb = std::atol(line + 6);
      std::fclose(f);
    }
    std::printf(" Peak host memory: %.3f MB\n Device memory in use: %.3f MB of %.3f MB\n", peak_kb / 1024.0,
                (tot - fr) / 1048576.0, tot / 1048576.0);
    std::fflush(stdout);
  });
}

// negf_get_version (negf_version, libnegf.F90:1295-1303)
void negf_get_version(int *handler) {
  api(handler, [&](Session &) {
    std::printf("(libnegf) version: 1.%20s  %10s\n", "hipnegf", __DATE__);
    std::fflush(stdout);
  });
}

void negf_set_params(int *handler, struct lnparams *params) {
  api(handler, [&](Session &s) { s.p = *params; });
}

void negf_get_params(int *handler, struct lnparams *params) {
  api(handler, [&](Session &s) { *params = s.p; });
}

void negf_create_hs(int *handler, int nHk) {
  api(handler, [&](Session &s) {
    if (nHk < 1) throw std::runtime_error("nHk < 1");
    s.H.assign(nHk, Csr());
    ++s.gen;
    s.S.assign(nHk, Csr());
    s.s_id.assign(nHk, 0);
  });
}

static void set_mat(Session &s, bool is_h, int nrow, const double2 *A, const int *JA,
                    const int *IA, int iK) {
  const int k = iK - 1;
  if (k < 0 || k >= (int)s.H.size())
    throw std::runtime_error("Error: set_H with index > allocated array. Call create_HS with correct size");
  if (is_h) s.H[k] = make_csr(nrow, A, JA, IA);
  else {
    s.S[k] = make_csr(nrow, A, JA, IA);
    s.s_id[k] = 0;
  }
  s.cur = k;
  ++s.gen;
}

void negf_set_h(int *handler, int nrow, const double *A, int *JA, int *IA, int iK) {
  api(handler, [&](Session &s) { set_mat(s, true, nrow, (const double2 *)A, JA, IA, iK); });
}
void negf_set_s(int *handler, int nrow, const double *A, int *JA, int *IA, int iK) {
  api(handler, [&](Session &s) { set_mat(s, false, nrow, (const double2 *)A, JA, IA, iK); });
}

static void set_cp(Session &s, bool is_h, int nrow, const double *re, const double *im,
                   const int *ja, const int *ia) {
  const int nnz = ia[nrow] - ia[0];
  std::vector<double2> A(nnz);
  for (int i = 0; i < nnz; ++i) A[i] = make_double2(re[i], im[i]);
  set_mat(s, is_h, nrow, A.data(), ja, ia, 1);
}
void negf_set_h_cp(int *handler, int nrow, double *a_re, double *a_im, int *ja, int *ia) {
  api(handler, [&](Session &s) { set_cp(s, true, nrow, a_re, a_im, ja, ia); });
}
void negf_set_s_cp(int *handler, int nrow, double *a_re, double *a_im, int *ja, int *ia) {
  api(handler, [&](Session &s) { set_cp(s, false, nrow, a_re, a_im, ja, ia); });
}

void negf_set_s_id(int *handler, int nrow, int iK) {
  api(handler, [&](Session &s) {
    const int k = iK - 1;
    if (k < 0 || k >= (int)s.S.size())
      throw std::runtime_error("Error: set_S_id with index > allocated array. Call create_HS with correct size");
    s.S[k] = Csr();
    s.S[k].n = 0;
    s.s_id[k] = 1;
    ++s.gen;
    (void)nrow;
  });
}

void negf_set_kpoint(int *handler, int kpoint) {
  api(handler, [&](Session &s) {
    s.p.kpoint = kpoint;
    if (kpoint >= 1 && kpoint <= (int)s.H.size() && s.cur != kpoint - 1) {
      s.cur = kpoint - 1;
      ++s.gen;
    }
  });
}

void negf_solve_landauer(int *handler) {
  api(handler, [](Session &s) { compute_current(s); });
}

// compute_ldos (libnegf.F90:1893-1901; the reference's C API has no binding)
void negf_compute_ldos(int *handler) {
  api(handler, [](Session &s) { compute_ldos(s); });
}

void negf_calculate_dephasing_transmission(int *handler) {
  api(handler, [](Session &s) { compute_dephasing_transmission(s); });
}

void negf_current(int *handler, double *current, char *unitsOfH, char *unitsOfJ) {
  api(handler, [&](Session &s) {
    compute_current(s);
    *current = (s.currents.empty() ? 0.0 : s.currents[0]) * convert_current(cstr(unitsOfH), cstr(unitsOfJ));
  });
}

void negf_set_elph_dephasing(int *handler, double *coupling, int coupling_size, int scba_niter) {
  api(handler, [&](Session &s) {
    s.deph = true;
    s.coupling.assign(coupling, coupling + coupling_size);
    s.scba_niter = scba_niter;
  });
}

void negf_set_scba_tolerances(int *handler, double elastic_tol, double inelastic_tol) {
  api(handler, [&](Session &s) {
    s.p.scba_elastic_tol = elastic_tol;
    s.p.scba_inelastic_tol = inelastic_tol;
  });
}

void negf_get_energies(int *handler, int *npoints, double *re_en, double *im_en, int copy) {
  api(handler, [&](Session &s) {
    *npoints = (int)s.en.size();
    if (copy)
      for (int i = 0; i < *npoints; ++i) {
        re_en[i] = s.en[i];
        im_en[i] = s.en_im.size() == s.en.size() ? s.en_im[i] : 0.0;
      }
  });
}

void negf_get_currents(int *handler, int *npairs, double *currents, int copy) {
  api(handler, [&](Session &s) {
    *npairs = (int)s.currents.size();
    if (copy)
      for (int i = 0; i < *npairs; ++i) currents[i] = s.currents[i];
  });
}

void negf_associate_transmission(int *handler, int tr_shape[2], double **tr_pointer) {
  api(handler, [&](Session &s) {
    tr_shape[0] = s.npoints;
    tr_shape[1] = s.tunn.empty() ? 0 : s.npairs;
    *tr_pointer = s.tunn.empty() ? nullptr : s.tunn.data();
  });
}

void negf_associate_energy_current(int *handler, int currents_shape[2], double **currents_pointer) {
  api(handler, [&](Session &s) {
    currents_shape[0] = s.npoints;
    currents_shape[1] = s.curr_mat.empty() ? 0 : s.ncont;
    *currents_pointer = s.curr_mat.empty() ? nullptr : s.curr_mat.data();
  });
}

// read_HS (libnegf.F90:202-254) with read_H (input_output.F90:41-264) for the
// formatted PETSc ASCII form, full matrix: two files (real, imaginary), each a
// 4-line header ("% Size = nrow ncol", "% Nonzeros = nnz", two more lines)
// and nnz lines "j i value" with 1-based indices.  target 0: H, 1: S.
void negf_read_hs(int *handler, const char *real_path, const char *imag_path, int target_matrix) {
  api(handler, [&](Session &s) {
    if (target_matrix != 0 && target_matrix != 1)
      throw std::runtime_error("libNEGF error. Wrong target_matrix: must be 0 (H) or 1 (S)");
    struct Part { int nrow = 0, ncol = 0, nnz = 0; std::vector<int> i, j; std::vector<double> v; };
    auto read_part = [](const char *path) {
      FILE *f = std::fopen(path, "r");
      if (!f) throw std::runtime_error("libNEGF error. Matrix files not found");
      Part p;
      char line[4096];
      auto next = [&]() -> std::string {
        if (!std::fgets(line, sizeof(line), f)) throw std::runtime_error("matrix file ended early");
        return std::string(line);
      };
      auto fields = [](const std::string &l) {
        std::vector<std::string> out;
        std::string cur;
        for (char ch : l) {
          if (ch == ' ' || ch == '\t' || ch == ',' || ch == '\n' || ch == '\r') {
            if (!cur.empty()) out.push_back(cur);
            cur.clear();
          } else {
            cur += ch;
          }
        }
        if (!cur.empty()) out.push_back(cur);
        return out;
      };
      try {
        std::vector<std::string> h1 = fields(next()), h2 = fields(next());
        if (h1.size() < 5 || h2.size() < 4) throw std::runtime_error("bad PETSc header");
        p.nrow = std::stoi(h1[3]);
        p.ncol = std::stoi(h1[4]);
        p.nnz = std::stoi(h2[3]);
        next();
        next();
        for (int k = 0; k < p.nnz; ++k) {
          std::vector<std::string> t = fields(next());
          if (t.size() < 3) throw std::runtime_error("bad matrix entry line");
          p.j.push_back(std::stoi(t[0]));  // the file lists (j, i, value)
          p.i.push_back(std::stoi(t[1]));
          std::string v = t[2];
          for (char &ch : v) if (ch == 'd' || ch == 'D') ch = 'e';
          p.v.push_back(std::stod(v));
        }
      } catch (...) {
        std::fclose(f);
        throw;
      }
      std::fclose(f);
      return p;
    };
    const Part re = read_part(real_path), im = read_part(imag_path);
    if (re.nrow != im.nrow) throw std::runtime_error("nrow Error");
    if (re.nnz != im.nnz) throw std::runtime_error("nnz Error");
    if (re.nnz == 0) throw std::runtime_error("nnz Error");
    for (int k = 0; k < re.nnz; ++k)
      if (re.i[k] != im.i[k] || re.j[k] != im.j[k]) throw std::runtime_error("Index Error");
    // coo2csr: row-major order, entries of a row in file order
    const int n = re.nrow;
    std::vector<int> ia(n + 1, 0), ja(re.nnz);
    std::vector<double2> a(re.nnz);
    for (int k = 0; k < re.nnz; ++k) {
      if (re.i[k] < 1 || re.i[k] > n || re.j[k] < 1 || re.j[k] > n)
        throw std::runtime_error("matrix index out of range");
      ++ia[re.i[k]];
    }
    for (int r = 0; r < n; ++r) ia[r + 1] += ia[r];
    std::vector<int> fill(ia.begin(), ia.end() - 1);
    for (int k = 0; k < re.nnz; ++k) {
      const int pos = fill[re.i[k] - 1]++;
      ja[pos] = re.j[k] - 1;
      a[pos] = make_double2(re.v[k], im.v[k]);
    }
    // create_HS(negf, 1): one k slot
    if (s.H.size() != 1) {
      s.H.assign(1, Csr());
      s.S.assign(1, Csr());
      s.s_id.assign(1, 0);
    }
    set_mat(s, target_matrix == 0, n, a.data(), ja.data(), ia.data(), 1);
  });
}

// set_outpath (libnegf.F90:943-949)
// set_scratch, libnegf.F90:922-940 (the folder must exist, as in the reference)
void negf_set_scratch(int *handler, const char *scratch_path) {
  api(handler, [&](Session &s) { s.scratch = cstr(scratch_path) + "/"; });
}

void negf_set_output(int *handler, const char *out_path) {
  api(handler, [&](Session &s) { s.out_path = cstr(out_path) + "/"; });
}

// write_tunneling_and_dos (libnegf.F90:2041-2099): tunneling_<kpoint>_<rank>.dat
// and localDOS_<kpoint>_<rank>.dat in the output path, Fortran E17.8 fields
void negf_write_tunneling_and_dos(int *handler) {
  api(handler, [&](Session &s) {
    auto e17 = [](double v) {
      // E17.8: 0.dddddddd mantissa, 2-digit exponent (3 digits without the E)
      char b[64], out[64];
      if (v == 0.0 || !std::isfinite(v)) {
        if (v == 0.0) std::snprintf(out, sizeof(out), "%17s", "0.00000000E+00");
        else std::snprintf(out, sizeof(out), "%17s", std::isnan(v) ? "NaN" : (v > 0 ? "Infinity" : "-Infinity"));
        return std::string(out);
      }
      std::snprintf(b, sizeof(b), "%.7e", std::fabs(v));  // d.ddddddde+XX
      const int ex = std::atoi(b + 10) + 1;
      char mant[16] = {'0', '.', b[0], b[2], b[3], b[4], b[5], b[6], b[7], b[8], 0};
      char body[48];
      if (std::abs(ex) <= 99) std::snprintf(body, sizeof(body), "%s%sE%c%02d", v < 0 ? "-" : "", mant, ex < 0 ? '-' : '+', std::abs(ex));
      else std::snprintf(body, sizeof(body), "%s%s%c%03d", v < 0 ? "-" : "", mant, ex < 0 ? '-' : '+', std::abs(ex));
      std::snprintf(out, sizeof(out), "%17s", body);
      return std::string(out);
    };
    char kp[16];
    std::snprintf(kp, sizeof(kp), "%06d", s.p.kpoint);
    const std::string tag = std::string(kp) + "_000000.dat";
    if (!s.tunn.empty()) {
      const std::string fn = s.out_path + "tunneling_" + tag;
      FILE *f = std::fopen(fn.c_str(), "w");
      if (!f) throw std::runtime_error("cannot open " + fn);
      const int n = s.npoints;
      for (int i = 0; i < n; ++i) {
        const double E = s.p.emin + s.p.estep * i;
        std::string l = e17(E * s.p.eneconv);
        for (int ip = 0; ip < s.npairs; ++ip) l += e17(s.tunn[(size_t)ip * n + i]);
        std::fprintf(f, "%s\n", l.c_str());
      }
      std::fclose(f);
    }
    if (!s.ldos.empty() && s.ndos > 0) {
      const std::string fn = s.out_path + "localDOS_" + tag;
      FILE *f = std::fopen(fn.c_str(), "w");
      if (!f) throw std::runtime_error("cannot open " + fn);
      const int n = s.npoints;
      for (int i = 0; i < n; ++i) {
        const double E = s.p.emin + s.p.estep * i;
        std::string l = e17(E * s.p.eneconv);
        for (int l2 = 0; l2 < s.ldos_cols; ++l2) l += e17(s.ldos[(size_t)l2 * n + i] / s.p.eneconv);
        std::fprintf(f, "%s\n", l.c_str());
      }
      std::fclose(f);
    }
  });
}

// compute_density_quasiEq (libnegf.F90:1801-1830) = quasiEq_int_n / _p
// (integrations.F90:3002-3258): for orbital k, q_k = Re i (t_k - conj t_k)
// with t_k = sum_i ww_i(k) G^r_kk(z_i) over a horizontal segment at height
// Lambda spanning every orbital's band window (Np(2) points), a vertical
// segment at the window's low (electrons) or high (holes) end (Np(1) points)
// and the Fermi poles of the orbital's own quasi-Fermi level.  The weights
// depend on the orbital only through that level, so orbitals sharing a level
// share one batched pass over the points; rho is not kept.
void compute_density_quasi(Session &s, int particle, double *q, int ndofs, const double *Ec,
                           const double *Ev, const double *mun, const double *mup) {
  if (particle != 1 && particle != -1)
    throw std::runtime_error("libNEGF error. In compute_density_quasiEq, unknown particle");
  if (s.ncont < 1) throw std::runtime_error("density needs at least one contact");
  hn::DevStruct *st = &device_struct(s);
  hn::Engine &eng = engine(s);
  const int ncont = s.ncont, N = s.H[s.cur].n;
  if (ndofs < 1 || ndofs > N) throw std::runtime_error("compute_density_quasiEq: ndofs outside 1..size of H");
  const cplx I(0.0, 1.0);
  double kbT = s.p.kbt_dm[0];
  for (int c = 1; c < ncont; ++c) kbT = std::max(kbT, s.p.kbt_dm[c]);
  const double Lambda = s.p.n_poles == 0 ? 0.5 * kbT * kPi : 2.0 * s.p.n_poles * kbT * kPi;
  const double Omega = s.p.n_kt * kbT;
  const double *lev = particle == 1 ? mun : mup;
  double lo = 1e300, hi = -1e300;  // min(minE), max(maxE)
  for (int k = 0; k < ndofs; ++k) {
    lo = std::min(lo, particle == 1 ? Ec[k] - s.p.deltaec : Ev[k] - Omega);
    hi = std::max(hi, particle == 1 ? Ec[k] + Omega : Ev[k] + s.p.deltaev);
  }
  // the two segments: (start, end, points) in the order of the reference
  const int *np = particle == 1 ? s.p.np_n : s.p.np_p;
  struct Seg { cplx a, b; int n; };
  std::vector<Seg> sg;
  if (particle == 1) {
    sg.push_back({lo + I * Lambda, hi + I * Lambda, np[1]});
    sg.push_back({cplx(lo, 0.0), lo + I * Lambda, np[0]});
  } else {
    sg.push_back({hi + I * Lambda, lo + I * Lambda, np[1]});
    sg.push_back({cplx(hi, 0.0), hi + I * Lambda, np[0]});
  }
  std::vector<cplx> zs, base;  // segment points and their level-free weight factor
  for (const Seg &g : sg) {
    std::vector<double> x, wt;
    gauleg(0.0, 1.0, x, wt, g.n);
    const cplx zd = g.b - g.a;
    for (size_t i = 0; i < x.size(); ++i) {
      zs.push_back(g.a + x[i] * zd);
      base.push_back((double)particle * s.p.g_spin * s.p.kwght * wt[i] * zd / (2.0 * kPi));
    }
  }
  std::map<double, std::vector<int>> groups;
  for (int k = 0; k < ndofs; ++k) groups[lev[k]].push_back(k);
  hn::SolveOpts o = dm_opts(s);
  const std::vector<long long> lay = hn::Engine::density_layout(*st);
  const long long tot = lay.back();
  s.item_status.clear();
  for (const auto &gr : groups) {
    const double mu = gr.first;
    std::vector<cplx> z = zs, w(zs.size());
    for (size_t i = 0; i < zs.size(); ++i)
      w[i] = base[i] * (particle == 1 ? fermi_c(zs[i], mu, kbT) : fermi_c(-zs[i], -mu, kbT));
    for (int i = 1; i <= s.p.n_poles; ++i) {
      z.push_back(mu + I * kbT * kPi * (2.0 * i - 1.0));
      w.push_back(-(double)particle * I * kbT * s.p.g_spin);
    }
    if (z.empty()) {
      for (int k : gr.second) q[k] = 0.0;
      continue;
    }
    SparseAcc T = place_blocks(s, *st, density_run(s, *st, eng, o, tot, s.ncont + 1, 0, z, w, {}), lay, N,
                               s.outer == 2);
    for (int k : gr.second) {
      cplx t(0.0, 0.0);
      for (const auto &e : T.rows[k])
        if (e.first == k) t += cplx(e.second.x, e.second.y);
      q[k] = -2.0 * t.imag();  // Re i (t - conj t)
    }
  }
}

// negf_density_quasi_equilibrium (libnegf_api.f90:753-775): all by reference
void negf_density_quasi_equilibrium(int *handler, int *ndofs, double *density, int *particle, double *Ec,
                                    double *Ev, double *mu_n, double *mu_p) {
  api(handler, [&](Session &s) { compute_density_quasi(s, *particle, density, *ndofs, Ec, Ev, mu_n, mu_p); });
}

// negf_density_efa (libnegf_api.f90:724-749): all arguments by reference
void negf_density_efa(int *handler, int *ndofs, double *density, int *particle, bool *contacts) {
  api(handler, [&](Session &s) {
    if (*contacts)
      throw std::runtime_error("negf_density_efa: the bulk contact density (contacts = true) is not provided");
    compute_density_efa(s, *particle, density, *ndofs);
  });
}

// compute_density_dft (libnegf.F90:1664-1679)
void negf_solve_density_dft(int *handler) {
  api(handler, [](Session &s) { compute_density(s); });
}

// rho as 1-based CSR (libnegf_api.f90:836-861)
void negf_get_dm(int *handler, int *nnz, int *nrow, int *rowpnt, int *colind, double *re_nzval,
                 double *im_nzval, int copy) {
  api(handler, [&](Session &s) {
    *nnz = (int)s.rho_col.size();
    *nrow = s.rho_n;
    if (!copy) return;
    for (int r = 0; r <= s.rho_n; ++r) rowpnt[r] = s.rho_ptr[r] + 1;
    for (size_t k = 0; k < s.rho_col.size(); ++k) {
      colind[k] = s.rho_col[k] + 1;
      re_nzval[k] = s.rho_val[k].x;
      im_nzval[k] = s.rho_val[k].y;
    }
  });
}

// negf_set_outer (libnegf_api.f90:1052-1065)
void negf_set_outer(int *handler, int *outer) {
  api(handler, [&](Session &s) {
    if (*outer < 0 || *outer > 2) throw std::runtime_error("outer must be 0, 1 or 2");
    s.outer = *outer;
  });
}

// init_ldos (libnegf.F90:971-988)
void negf_init_ldos(int *handler, int nldos) {
  api(handler, [&](Session &s) {
    if (nldos < 0) throw std::runtime_error("nldos < 0");
    s.ndos = nldos;
    s.dos_proj.assign(nldos, std::vector<int>());
  });
}

// set_ldos_intervals (libnegf.F90:1010-1024): interval ii is istart..iend for
// ii = 1..ndos_proj (the count given to init_ldos)
void negf_set_ldos_intervals(int *handler, int nldos, int *istart, int *iend) {
  api(handler, [&](Session &s) {
    (void)nldos;
    for (int ii = 0; ii < s.ndos; ++ii) {
      std::vector<int> v;
      for (int jj = 0; jj < iend[ii] - istart[ii] + 1; ++jj) v.push_back(istart[ii] + jj);
      s.dos_proj[ii] = v;
    }
  });
}

// set_ldos_indexes (libnegf.F90:1031-1041), ildos 1-based
void negf_set_ldos_indexes(int *handler, int ildos, int idx_size, int *idx) {
  api(handler, [&](Session &s) {
    if (ildos < 1 || ildos > s.ndos) throw std::runtime_error("ildos out of range (negf_init_ldos)");
    s.dos_proj[ildos - 1].assign(idx, idx + std::max(idx_size, 0));
  });
}

// associate_ldos (libnegf.F90:1416-1426): ldos_mat, shape (npoints, ndos_proj)
void negf_associate_ldos(int *handler, int ldos_shape[2], double **ldos_pointer) {
  api(handler, [&](Session &s) {
    const bool have = !s.ldos.empty();
    ldos_shape[0] = have ? s.npoints : 0;
    ldos_shape[1] = have ? s.ldos_cols : 0;
    *ldos_pointer = have ? s.ldos.data() : nullptr;
  });
}

// set_tun_indexes (libnegf.F90:1044-1058); n < 0 removes the projection
void negf_set_tun_indexes(int *handler, int n, int *idx) {
  api(handler, [&](Session &s) {
    s.tun_set = n >= 0;
    s.tun_proj.assign(idx, idx + std::max(n, 0));
  });
}

void negf_set_precision(int *handler, int precision) {
  api(handler, [&](Session &s) { s.precision = precision; });
}
void negf_set_batch(int *handler, int max_batch) {
  api(handler, [&](Session &s) { s.max_batch = max_batch; });
}
int negf_get_status(int *handler) {
  Session *s = sess(handler);
  return s ? s->status : -1;
}
const char *negf_get_last_error(int *handler) {
  Session *s = sess(handler);
  return s ? s->err.c_str() : "invalid handler";
}
void negf_get_item_status(int *handler, int *npoints, int *status, int *ncyc, int copy) {
  api(handler, [&](Session &s) {
    *npoints = (int)s.item_status.size();
    if (!copy) return;
    for (int i = 0; i < *npoints; ++i) status[i] = s.item_status[i];
    if (ncyc)
      for (size_t i = 0; i < s.item_ncyc.size(); ++i) ncyc[i] = s.item_ncyc[i];
  });
}

}  // extern "C"
