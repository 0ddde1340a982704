// hn_engine.h -- batched per-(E,k) Green's-function engine (host side).
#pragma once
#include <functional>
#include <memory>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "hn_common.h"
#include "hn_inv.h"
#include "hn_kernels.h"

namespace hn {

// Block kinds of a structure slot (device copy of one H(k), S(k)).
enum BlockKind : int {
  BK_HD = 0, BK_SD, BK_HU, BK_SU, BK_HL, BK_SL,  // device: diag, upper (i,i+1), lower (i+1,i)
  BK_H00, BK_S00, BK_H01, BK_S01, BK_HMC, BK_SMC, // contacts
  // contact surface block (ns orbitals before the two PLs): surface, surface
  // -> PL1 and PL1 -> surface; HMC/SMC then span surface + PL1
  BK_HSS, BK_SSS, BK_HS1, BK_SS1, BK_H1S, BK_S1S,
  BK_COUNT
};

// S modes
enum SMode : int { S_DIAG = 1, S_OFFDIAG = 2, S_CONTACT = 4 };

// Device-resident block-tridiagonal structure with `nslots` slots of identical
// layout (one slot per transverse k or per request variant).
struct DevStruct {
  int nbl = 0, ncont = 0, smode = 0, nslots = 0;
  std::vector<int> bsz, cblk, nc;
  std::vector<int> ns;  // contact surface block sizes (0: none)
  // order of the contact's surface Green's function: surface + one PL
  int ngs(int c) const { return nc[c] + ns[c]; }
  std::vector<long long> off[BK_COUNT];
  long long slot_elems = 0;
  double2 *pool = nullptr;
  int device = 0;
  int mirror_state = -1;  // contacts 1/2 mirror images in every slot: -1 unknown, 0 no, 1 yes
  unsigned long long version = 0;  // bumped whenever block contents change (cache keys)
  unsigned long long uid = 0;      // process-unique id of this structure
  // fictitious contacts (contselfenergy.F90:175-185): GS = -i pi dos I, no decimation
  int fict[kMaxCont] = {0};
  double contact_dos[kMaxCont] = {0.0};
  void touched() { mirror_state = -1; ++version; }
  // tun_proj (libnegf.F90:1044-1058) as a device row mask over the device
  // orbitals: only rows with a nonzero byte enter transmission traces
  // (get_tun_mask, iterative_cpu.F90:585-617).  Null: every row.
  unsigned char *tun_mask = nullptr;
  // host 1-based orbital indexes; n == 0 clears the mask
  void set_tun_proj(int n, const int *idx);
  const unsigned char *tun_rows(int blk) const {
    return tun_mask ? tun_mask + row_off(blk) : nullptr;
  }
  long long row_off(int blk) const {
    long long o = 0;
    for (int i = 0; i < blk; ++i) o += bsz[i];
    return o;
  }
  int ndev() const { return (int)row_off(nbl); }

  bool owned = true;       // pool allocated (and freed) by this object

  DevStruct(int nbl, const int *bsz, int ncont, const int *cblk, const int *nc, int smode,
            int nslots, const int *ns = nullptr);
  // Same layout on caller-provided memory of at least layout_bytes(...) bytes,
  // zeroed on `zs` (stream-ordered; not freed by the destructor).
  DevStruct(int nbl, const int *bsz, int ncont, const int *cblk, const int *nc, int smode,
            int nslots, void *external, hipStream_t zs, const int *ns = nullptr);
  ~DevStruct();
  static size_t layout_bytes(int nbl, const int *bsz, int ncont, const int *cblk, const int *nc,
                             int smode, int nslots, const int *ns = nullptr);
  long long block_elems(int kind, int idx) const;
  BlkRef ref(int kind, int idx) const {
    BlkRef r;
    r.pool = pool;
    r.slot_stride = slot_elems;
    r.off = off[kind].empty() ? -1 : off[kind][idx];
    return r;
  }
  size_t bytes() const { return (size_t)slot_elems * nslots * sizeof(double2); }
  int nmax() const;
  bool uniform() const;
};

enum Precision : int { PREC_FP64 = 0, PREC_MIXED = 1 };

constexpr int kFusedMaxN = 16;  // largest block of the fused transmission kernel

// Device arena for engine workspaces (grown on demand, never shrunk).
struct Arena {
  char *base = nullptr;
  size_t cap = 0, used = 0;
  void reserve(size_t bytes);
  void reset() { used = 0; }
  template <typename T> T *take(size_t count) {
    size_t b = (count * sizeof(T) + 255) & ~(size_t)255;
    if (used + b > cap) throw std::runtime_error("arena overflow");
    T *p = (T *)(base + used);
    used += b;
    return p;
  }
  ~Arena();
};

struct SolveOpts {
  double delta = 1e-4;       // eta on the contacts (lib_param.F90:628 default)
  int ni = 1, nf = 2;        // 1-based contact pair
  int precision = PREC_FP64;
  int max_batch = 0;         // 0: automatic (memory budget)
  double sgfacc = 1e-9;      // SGFACC, ln_precision.F90:40
  int dec_maxit = 300;       // contselfenergy.F90:289
  double eps_drop = 0.0;     // EPS (ln_precision), default 0
  int ns_steps = 1;          // mixed precision: FP64 Newton-Schulz steps per block inverse
  int polish = 0;            // reserved
  int dec_crit = 0;          // 0: max|C1| (CPU path), 1: sum|Re|+|Im| (GPU path)
  int share_mirror = 1;      // one decimation for mirror-image contact pairs
  int lanes = 0;             // item groups solved concurrently, each on its own stream;
                             // 0: automatic (2 for >= 256 items of blocks >= 128, else 1)
  int graph = 0;             // replay the RGF sweep of a batch from a captured hipGraph
  int fused = 1;             // small uniform blocks: whole solves in one launch (hn_fused.hip)
  // surface Green's function cache (surface_green + ln_cache, contselfenergy.F90:138-173,
  // 251-259): 0 compute only; 1 load if cached, else compute and store; 2 compute and store
  int sgf_mode = 0;
  // device [nitems] imaginary parts added to delta on the contacts (complex
  // energies E + i (delta + Ei), e.g. contour points); null: none
  const double *dEi = nullptr;
};

// Local DOS alongside the transmission (calculate_transmissions_and_dos,
// iterative.F90:1821-1952).
struct DosOpts {
  double dos_delta = 0.0;     // extra imaginary part of the device energy (negf%dos_delta)
  int nldos = 0;              // number of projections
  int full = 0;               // one output per device orbital instead of projections
  const int *ptr = nullptr;   // device [nldos+1]: list l is idx[ptr[l] .. ptr[l+1]-1]
  const int *idx = nullptr;   // device, 1-based device orbitals
};

// Density-matrix integrand (contour_int / real_axis_int, integrations.F90:844-1176):
// acc += sum over items of w_b X_b for the blocks of one integration path.
struct DensityOpts {
  int mode = 0;                 // 0: G^r, contacts and device at E + i Ei (contour points)
                                // 1: G^n, contacts at E + i delta, device at E (real axis)
  const double2 *w = nullptr;   // device [nitems] weights
  const double *frm = nullptr;  // mode 1: device [nitems * (ncont+1)] occupations
  int ref = 0;                  // mode 1: 1-based reference contact (ncont + 1: none)
  int outer = 2;                // 0: device blocks, 1: + device/contact, 2: + contact/device
  int *status = nullptr;        // optional device [nitems]: per-item status words
  int real_device = 0;          // mode 0: the device at real E, only the contacts at E + i Ei
                                // (the G^r returned beside G^n, iterative.F90:263-276)
  // mode 0, optional: per item, gs_scale * Im diag(G^r S) of every device orbital
  // (G^r on the pattern of the device + contact S, outer >= 1 adding G^r_DC S_CD;
  // ldos_int, integrations.F90:241-246) ADDED into gs_diag[item * gs_ld + orbital]
  // (device, zeroed by the caller).  With gs_diag set, acc and w may be null.
  double *gs_diag = nullptr;
  long long gs_ld = 0;
  double gs_scale = 0.0;
};

struct SCBAOpts {
  const double *coupling = nullptr;  // device, per device orbital (squared inside, elphdd.F90:94)
  int niter = 0;                     // scba_niter (loop runs while iter <= niter)
  double tol = 1e-7;                 // scba_elastic_tol (lib_param.F90:653)
  const double *frm = nullptr;       // device, [ncont+1] occupations (frm[ref-1]: reference)
  int frm_per_item = 0;              // 1: frm holds nitems x (ncont+1) values
  int ref = 0;                       // 1-based reference contact (ncont+1: none)
};

// Inelastic electron-phonon model (ElPhonInel: elphinel.F90, self_energy.F90)
// in meir_wingreen's outer SCBA loop (integrations.F90:1804-1966).  Items are
// the (E,k) points of a uniform energy grid: item = iK * ne + iE.
struct InelOpts {
  int ne = 0, nk = 0;
  const double *kw = nullptr;  // host [nk] k weights of the curr_mat sum
  const double *K = nullptr;   // host [ndz * nk * nk], K[iz + ndz (iK + nk iQ)]
  int ndz = 0;
  const int *iz = nullptr;     // host [ndev] z bin of each device orbital
  int shift = 0;               // nint(wq / dE): energy points per phonon quantum
  double Nq = 0.0;             // Bose occupation of the phonon mode
  int tridiag = 0;             // with the (i,i+1) / (i+1,i) self-energy blocks
  int niter = 0;               // outer passes run while p <= niter and not converged
  double tol = 1e-7;           // max |J - J_old| over contacts (check_J_convergence)
  int integration = 0;         // integrate_el_meir_wingreen rule (0 trapz, 1 simpson13, 2 simpson38)
  double estep = 0.0, g_spin = 2.0;
};

// Block stores of one inelastic solve, [block][item][n*n] with item stride n*n.
struct InelIO {
  int i0 = 0, NI = 0, tridiag = 0;
  double2 *G[6] = {};        // out: Grd, Gru, Grl, Gnd, Gnu, Gnl of each item's last pass
  const double2 *S[5] = {};  // in: Srd, Sru, Srl, Snd, Snu (null: none, outer pass 0)
  double2 *sigc = nullptr;   // contact self-energies [c][item][n*n] kept across passes
  int *stc = nullptr;        // their status [c][item]
  bool sigc_valid = false;
};

// integrate_el_meir_wingreen (integrations.F90:2774-2832)
double integrate_mw(int integration, const double *Iv, int npts, double estep);

// Surface Green's function cache (the role of ln_cache's TMatrixCacheMem for
// surface_green, contselfenergy.F90:138-173, 251-259): g_s of (structure and
// its contents version, slot, E, eta, contact) in one device buffer of
// fixed-size entries, least recently used out first.  One engine owns it, or
// several engines of one device share it (the serving workers): every use
// holds `mu`, and the device copies into or out of the pool are complete
// before it is released.
struct SgfCache {
  struct Key {
    unsigned long long uid, version, ebits, dbits;
    int slot, contact, nc;
    bool operator==(const Key &o) const {
      return uid == o.uid && version == o.version && ebits == o.ebits && dbits == o.dbits &&
             slot == o.slot && contact == o.contact && nc == o.nc;
    }
  };
  struct KeyHash {
    size_t operator()(const Key &k) const {
      size_t h = std::hash<unsigned long long>()(k.uid * 1000003ULL ^ k.version);
      h ^= std::hash<unsigned long long>()(k.ebits) + 0x9e3779b97f4a7c15ULL + (h << 6) + (h >> 2);
      h ^= std::hash<unsigned long long>()(k.dbits ^ ((unsigned long long)k.slot << 20) ^
                                          ((unsigned long long)k.contact << 8) ^ (unsigned long long)k.nc);
      return h;
    }
  };
  std::mutex mu;
  double2 *pool = nullptr;
  size_t bytes = 0;
  long long entry_elems = 0;
  int capacity = 0;
  std::unordered_map<Key, int, KeyHash> map;
  std::vector<Key> keys;
  std::vector<unsigned long long> used;  // LRU stamps
  std::vector<int> valid;
  std::vector<int> status;  // decimation status bits stored with each entry
  unsigned long long clock = 0;
  long long hits = 0, stores = 0;
  SgfCache(int device, size_t bytes);
  ~SgfCache();
  SgfCache(const SgfCache &) = delete;
  SgfCache &operator=(const SgfCache &) = delete;
  // entries of at least entry_elems elements (re-partitions, dropping every
  // entry, when a larger size is asked for); false: no capacity.  Caller holds mu.
  bool prepare(long long entry_elems);
  // the entry of k, inserted (evicting the least recently used) if absent.  Caller holds mu.
  int slot_for(const Key &k);
};

class Engine {
 public:
  explicit Engine(int device);
  ~Engine();
  hipStream_t stream() const { return stream_; }
  int device() const { return device_; }

  // Coherent transmission for nitems (E, slot) points.  All pointers device.
  // T[nitems], ncyc[nitems*ncont], status[nitems]; the pair is (o.ni, o.nf).
  void transmission(DevStruct &st, int nitems, const double *dE, const int *dslot,
                    const SolveOpts &o, double *dT, int *dncyc, int *dstatus);
  // One-time per-structure checks (mirror-image contacts) so that several
  // engines can then share the structure read-only.
  void prepare(DevStruct &st) { (void)mirror_contacts(st); }
  // device workspace bytes per item of a transmission batch on `st`
  static size_t transmission_item_bytes(const DevStruct &st, int npairs, const int *ni,
                                        const int *nf);
  // Same for npairs contact pairs (host arrays ni[], nf[], 1-based):
  // T[nitems*npairs] item-major.  One pair with two contacts and the lower
  // contact on block 1 takes the fast path, everything else the multi-terminal
  // column sweep (calculate_transmissions, iterative.F90:1669-1780).
  void transmission_pairs(DevStruct &st, int nitems, const double *dE, const int *dslot,
                          const SolveOpts &o, int npairs, const int *ni, const int *nf,
                          double *dT, int *dncyc, int *dstatus);

  // Transmission for npairs pairs and the local DOS per item, with every
  // Gr(i,i) block of the sweep: contacts at E + i delta, device at
  // E + i dos_delta.  T[nitems*npairs], ldos[nitems*nout] item-major
  // (nout = full ? ndev : nldos; ldos may be null when nout == 0).
  void transmission_dos(DevStruct &st, int nitems, const double *dE, const int *dslot,
                        const SolveOpts &o, int npairs, const int *ni, const int *nf,
                        const DosOpts &d, double *dT, double *dldos, int *dncyc, int *dstatus);
  static size_t dos_item_bytes(const DevStruct &st, int npairs, const int *ni, const int *nf);

  // Density-matrix integrand for nitems points, summed into acc (device,
  // packed column-major blocks in the order of density_layout: Gd[nbl],
  // Gu[nbl-1] = (i,i+1), Gl[nbl-1] = (i+1,i), Glc[ncont] (nd x nc),
  // Gcl[ncont] (nc x nd)).  dEi: mode 0 imaginary parts (device).
  void density(DevStruct &st, int nitems, const double *dE, const double *dEi, const int *dslot,
               const SolveOpts &o, const DensityOpts &d, double2 *acc);
  // offsets of the blocks in the accumulator; back() = total elements
  static std::vector<long long> density_layout(const DevStruct &st);

  // Elastic-dephasing SCBA + Meir-Wingreen current per item.
  // curr[nitems*ncont], iters[nitems], err[nitems], status[nitems];
  // gn_diag/gr_diag [nitems*ndev] optional (device).
  void scba_mw(DevStruct &st, int nitems, const double *dE, const int *dslot,
               const SolveOpts &o, const SCBAOpts &so, double *dcurr, int *diters, double *derr,
               int *dstatus, double2 *gn_diag, double2 *gr_diag);

  // meir_wingreen with the inelastic model: E[ne] (host, uniform grid),
  // frm[ne * (ncont+1)] (host) occupations per energy, so.coupling the elastic
  // dephasing (device, may be null: none), so.niter / so.tol its inner loop.
  // Out (host): curr_mat[ne * ncont] (iE + ne c, k-weighted), currents[ncont],
  // passes, err (last max |dJ|), status / iters [ne * nk] (contacts | inverses
  // over all passes; inner passes of the last outer pass).
  void inel_mw(DevStruct &st, const double *E, const double *frm, const SolveOpts &o,
               const SCBAOpts &so, const InelOpts &io, double *curr_mat, double *currents,
               int *passes, double *err, int *status, int *iters);

  // Single decimation with host matrices (cu_Zdecimation semantics).
  int decimation_host(int n, double2 *hGo, const double2 *hAo, const double2 *hBo,
                      const double2 *hCo, double acc, int crit, int *ncyc);

  int last_batch() const { return last_batch_; }
  // device memory for the surface Green's function cache (0 disables it)
  void set_sgf_cache(size_t bytes);
  // use a cache shared with other engines of the same device (null: none)
  void attach_sgf_cache(std::shared_ptr<SgfCache> c);
  const std::shared_ptr<SgfCache> &sgf_cache() const { return sgf_; }
  long long sgf_hits() const;
  long long sgf_stores() const;
  size_t sgf_cache_bytes() const;
  // One cached g_s (column-major, order ngs(c)) of item (slot, E) under the
  // key a solve with options o uses: copied to host memory (false when it is
  // not cached) or inserted from it, with its decimation status.  These serve
  // the C API's disk cache of surface functions.
  bool sgf_fetch(DevStruct &st, const SolveOpts &o, int slot, double E, int c, double2 *host,
                 int *status);
  void sgf_insert(DevStruct &st, const SolveOpts &o, int slot, double E, int c,
                  const double2 *host, int status);
  // Sub-engines (own stream and workspace) for concurrent item groups.
  Engine &lane(int i);
  size_t workspace_bytes() const { return arena_.cap; }

 private:
  int device_;
  hipStream_t stream_;
  Arena arena_;
  int *h_count_ = nullptr;  // pinned
  int last_batch_ = 0;
  std::vector<std::unique_ptr<Engine>> lanes_;
  // surface Green's function cache (own or shared with other engines)
  std::shared_ptr<SgfCache> sgf_;
  SgfCache::Key sgf_key(const DevStruct &st, double E, double delta, int slot, int c,
                        bool shared) const;
  bool sgf_shared(DevStruct &st, const SolveOpts &o);
  static long long sgf_entry_of(const DevStruct &st);  // elements per entry: max ngs^2
  struct GraphEntry {
    std::vector<long long> key;
    hipGraphExec_t exec = nullptr;
  };
  std::vector<GraphEntry> graphs_;  // captured sweeps, most recent last
  void clear_graphs();

  int pick_batch(const DevStruct &st, int nitems, const SolveOpts &o, size_t per_item) const;
  // whole two-contact transmission solves in one launch for blocks of at most
  // kFusedMaxN orbitals (hn_fused.hip); false when the call is not eligible
  bool t_fused(DevStruct &st, int nitems, const double *dE, const int *dslot, const SolveOpts &o,
               int ni, int nf, double *dT, int *dncyc, int *dstatus);
  // GSout (optional): the surface Green's functions, contact-major, stride sGS
  void contacts(DevStruct &st, int B, const double *dE, const int *dslot,
                const SolveOpts &o, double2 *Sig, long long sSig, int *ncyc_c, int *status_c,
                double2 *GSout = nullptr, long long sGS = 0);
  void decimate_z(int nc, int CB, double2 *Ao, double2 *Aos, double2 *Bo, double2 *Co,
                  double2 *Go, const SolveOpts &o, int *ncyc, int *status);
  bool mirror_contacts(DevStruct &st);
  void decimate_shared(int nc, int B, double2 *Ao, double2 *Aos0, double2 *Aos1, double2 *Bo,
                       double2 *Co, double2 *Go0, double2 *Go1, const SolveOpts &o, int *ncyc0,
                       int *ncyc1, int *status0, int *status1);
  void inverse(int n, double2 *A, long long sA, double2 *Ainv, long long sAinv, int batch,
               const int *active, int *status, int precision, const SolveOpts &o,
               int prof_items = -1);
  void t_batch(DevStruct &st, int B, const double *dE, const int *dslot,
               const SolveOpts &o, int npairs, const int *ni, const int *nf, double *dT,
               int *dncyc, int *dstatus);
  void t_sweep(DevStruct &st, int B, const double *dE, const int *dslot, const SolveOpts &o,
               int ni, int nf, const double2 *Sig, long long sSig, double *dT, int *dstatus);
  void t_sweep_mixed(DevStruct &st, int B, const double *dE, const int *dslot, const SolveOpts &o,
                     int ni, int nf, const double2 *Sig, long long sSig, double *dT, int *dstatus);
  void t_multi(DevStruct &st, int B, const double *dE, const int *dslot, const SolveOpts &o,
               int npairs, const int *ni, const int *nf, const double2 *Sig, long long sSig,
               double *dT, int *dstatus);
  void density_batch(DevStruct &st, int B, const double *dE, const double *dEi, const int *dslot,
                     const SolveOpts &o, const DensityOpts &d, const double2 *w, const double *frm,
                     double2 *acc, const std::vector<long long> &lay);
  void dos_batch(DevStruct &st, int B, const double *dE, const int *dslot, const SolveOpts &o,
                 int npairs, const int *ni, const int *nf, const DosOpts &d, double *dT,
                 double *dldos, int *dncyc, int *dstatus);
  // per-item ncyc (item-major) and status words from the contact stage's arrays
  void finish_items(int B, int ncont, const int *ncyc_c, const int *status_c, int *dncyc,
                    int *dstatus);
  void scba_batch(DevStruct &st, int B, const double *dE, const int *dslot,
                  const SolveOpts &o, const SCBAOpts &so, double *dcurr, int *diters,
                  double *derr, int *dstatus, double2 *gn_diag, double2 *gr_diag,
                  const InelIO *io = nullptr);
};

}  // namespace hn
