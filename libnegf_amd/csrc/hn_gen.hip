// hn_gen.hip -- on-device generation of the k-resolved random slab.
//
// Serving workloads carry only (seed, n, nbl, k): the H(k), S(k) blocks of a
// request are generated in HBM by a counter-based Gaussian generator instead
// of being built on the host and copied over PCIe (an n = 512, nbl = 200
// request is ~5 GB of blocks).  The model is the slab of
// libnegf_amd/workloads.py (KSlab): per-PL real-symmetric on-site blocks
// A0_p (sigma 0.3), couplings A1_p (sigma 0.1), transverse Bloch terms
// Bx, By (sigma 0.1), lead blocks A0c, A1c, and overlap S = I + s R with the
// same structure (sigma 1, scaled by s).  Values are not the numpy stream of
// KSlab; the distribution and structure are the same.
//
// Random numbers: element (i, j) of matrix `id` is a standard normal from
// the Box-Muller transform of two uniforms taken from SplitMix64 (Steele,
// Lea, Flood, OOPSLA 2014) applied to a key built from (seed, id, i, j).
#include <cmath>

#include "hn_gen.h"

namespace hn {

namespace {

constexpr int NT = 256;

__device__ inline unsigned long long splitmix(unsigned long long z) {
  z += 0x9e3779b97f4a7c15ULL;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

__device__ inline double gauss(unsigned long long seed, int id, int i, int j, int n) {
  const unsigned long long key = splitmix(seed * 0xd1342543de82ef95ULL + (unsigned long long)id);
  const unsigned long long h1 = splitmix(key + (unsigned long long)i * (unsigned long long)n + j);
  const unsigned long long h2 = splitmix(h1);
  const double u1 = ((double)(h1 >> 11) + 0.5) * 0x1.0p-53;
  const double u2 = (double)(h2 >> 11) * 0x1.0p-53;
  return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

struct GenArgs {
  double2 *base;          // first block of this kind in the slot
  long long stride;       // elements between consecutive blocks of the kind
  int n;
  unsigned long long seed;
  int id0;                // matrix id of block b: id0 + b
  int sym;                // base matrix is (G + G^T) / sqrt 2
  int transpose;          // use the base matrix transposed
  double sigma;           // base scale
  int onsite;             // add the Bloch sum of (bx, by)
  int bx, by;
  double bsigma, kx, ky;
  double scale;           // overall factor
  int identity;           // add I
};

__global__ __launch_bounds__(NT) void k_gen(GenArgs a) {
  const long long e = (long long)blockIdx.x * NT + threadIdx.x;
  const int n = a.n;
  if (e >= (long long)n * n) return;
  const int b = blockIdx.y;
  int r = (int)(e % n), c = (int)(e / n);
  if (a.transpose) { const int t = r; r = c; c = t; }
  const int id = a.id0 + b;
  double v = gauss(a.seed, id, r, c, n);
  if (a.sym) v = (v + gauss(a.seed, id, c, r, n)) * 0.7071067811865476;
  double re = a.sigma * v, im = 0.0;
  if (a.onsite) {
    // Bx e^{ikx} + Bx^T e^{-ikx} + By e^{iky} + By^T e^{-iky}
    const double xp = gauss(a.seed, a.bx, r, c, n), xm = gauss(a.seed, a.bx, c, r, n);
    const double yp = gauss(a.seed, a.by, r, c, n), ym = gauss(a.seed, a.by, c, r, n);
    const double cx = cos(a.kx), sx = sin(a.kx), cy = cos(a.ky), sy = sin(a.ky);
    re += a.bsigma * ((xp + xm) * cx + (yp + ym) * cy);
    im += a.bsigma * ((xp - xm) * sx + (yp - ym) * sy);
  }
  re *= a.scale;
  im *= a.scale;
  if (a.identity && r == c) re += 1.0;
  a.base[(long long)b * a.stride + e] = make_double2(re, im);
}

// matrix ids
enum : int {
  ID_BX = 1, ID_BY, ID_A0C, ID_A1C, ID_RX, ID_RY, ID_R0C, ID_R1C,
  ID_A0 = 1 << 16, ID_A1 = 2 << 16, ID_R0 = 3 << 16, ID_R1 = 4 << 16
};

}  // namespace

double kslab_overlap_scale(int n) {
  // ||G||_2 ~ 2 sqrt(n) sigma for an n x n Gaussian (and for (G + G^T)/sqrt 2);
  // R(k) has on-site R0 + Bloch terms (2||Rx|| + 2||Ry||) and two couplings:
  // ~ 14 sqrt(n); 25 % margin.  s R then has norm <= 0.5, so S = I + s R > 0.
  const double bound = 1.25 * 14.0 * std::sqrt((double)n);
  return 0.05 * std::min(1.0, 10.0 / bound);
}

void gen_kslab(DevStruct &st, int slot, unsigned long long seed, double kx, double ky,
               int with_overlap, hipStream_t s) {
  const int nbl = st.nbl;
  if (!st.uniform() || st.ncont != 2 || st.cblk[0] != nbl || st.cblk[1] != 1 || st.ns[0] || st.ns[1])
    throw std::runtime_error("gen_kslab: structure must have equal PL/contact sizes, "
                             "contact 1 on the last block and contact 2 on block 1");
  if (with_overlap && st.smode != (S_DIAG | S_OFFDIAG | S_CONTACT))
    throw std::runtime_error("gen_kslab: overlap needs smode = 7");
  if (slot < 0 || slot >= st.nslots) throw std::runtime_error("gen_kslab: slot out of range");
  const int n = st.bsz[0];
  const double ss = kslab_overlap_scale(n);
  double2 *slot_base = st.pool + (long long)slot * st.slot_elems;
  auto launch = [&](int kind, int count, GenArgs a) {
    if (count <= 0) return;
    a.base = slot_base + st.off[kind][0];
    a.stride = count > 1 ? st.off[kind][1] - st.off[kind][0] : 0;
    a.n = n;
    a.seed = seed;
    a.kx = kx;
    a.ky = ky;
    const long long nn = (long long)n * n;
    hipLaunchKernelGGL(k_gen, dim3((unsigned)((nn + NT - 1) / NT), count), dim3(NT), 0, s, a);
    HN_CHECK(hipGetLastError());
  };
  auto one = [&](int kind, int idx, GenArgs a) {
    a.base = slot_base + st.off[kind][idx];
    a.stride = 0;
    a.n = n;
    a.seed = seed;
    a.kx = kx;
    a.ky = ky;
    const long long nn = (long long)n * n;
    hipLaunchKernelGGL(k_gen, dim3((unsigned)((nn + NT - 1) / NT), 1), dim3(NT), 0, s, a);
    HN_CHECK(hipGetLastError());
  };
  GenArgs hd{};
  hd.id0 = ID_A0; hd.sym = 1; hd.sigma = 0.3; hd.onsite = 1; hd.bx = ID_BX; hd.by = ID_BY;
  hd.bsigma = 0.1; hd.scale = 1.0;
  launch(BK_HD, nbl, hd);
  GenArgs hu{};
  hu.id0 = ID_A1; hu.sigma = 0.1; hu.scale = 1.0;
  launch(BK_HU, nbl - 1, hu);
  GenArgs hl = hu;
  hl.transpose = 1;
  launch(BK_HL, nbl - 1, hl);
  GenArgs h00 = hd;
  h00.id0 = ID_A0C;
  GenArgs a1c{};
  a1c.id0 = ID_A1C; a1c.sigma = 0.1; a1c.scale = 1.0;
  GenArgs a1ct = a1c;
  a1ct.transpose = 1;
  one(BK_H00, 0, h00);
  one(BK_H00, 1, h00);
  one(BK_H01, 0, a1c);   // contact 1 (on the last block): H01 = A1c, HMC = A1c
  one(BK_HMC, 0, a1c);
  one(BK_H01, 1, a1ct);  // contact 2 (on block 1): the mirror image
  one(BK_HMC, 1, a1ct);
  if (with_overlap) {
    GenArgs sd{};
    sd.id0 = ID_R0; sd.sym = 1; sd.sigma = 1.0; sd.onsite = 1; sd.bx = ID_RX; sd.by = ID_RY;
    sd.bsigma = 1.0; sd.scale = ss; sd.identity = 1;
    launch(BK_SD, nbl, sd);
    GenArgs su{};
    su.id0 = ID_R1; su.sigma = 1.0; su.scale = ss;
    launch(BK_SU, nbl - 1, su);
    GenArgs sl = su;
    sl.transpose = 1;
    launch(BK_SL, nbl - 1, sl);
    GenArgs s00 = sd;
    s00.id0 = ID_R0C;
    GenArgs r1c{};
    r1c.id0 = ID_R1C; r1c.sigma = 1.0; r1c.scale = ss;
    GenArgs r1ct = r1c;
    r1ct.transpose = 1;
    one(BK_S00, 0, s00);
    one(BK_S00, 1, s00);
    one(BK_S01, 0, r1c);
    one(BK_SMC, 0, r1c);
    one(BK_S01, 1, r1ct);
    one(BK_SMC, 1, r1ct);
  }
  st.touched();
}

}  // namespace hn
