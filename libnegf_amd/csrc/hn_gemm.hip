// hn_gemm.hip -- batched complex GEMM on CDNA4 matrix cores.
//
// Complex products use four real MFMAs:  Cr += Ar Br - Ai Bi,
// Ci += Ar Bi + Ai Br.  K is walked in steps of 16 complex elements staged
// through LDS as separate real/imaginary planes (k-major).  In the wide
// kernel the LDS stage is double-buffered: while the matrix cores work on
// step t from one buffer, step t+1 (prefetched into registers during step
// t-1) is stored into the other and step t+2 is fetched from memory -- one
// barrier per K step.  The narrow kernel (64 x 64, the small blocks) keeps
// one stage and fetches its C tile in the epilogue: 33 KiB of LDS, 126 VGPRs
// and 64 accumulation registers, two waves per SIMD (1.1x faster at n = 64
// than the double-buffered form, which held the same occupancy).
//
// FP64 uses v_mfma_f64_4x4x4 (four independent 4x4x4 blocks per
// instruction), which sustains 67 TF/s on MI355X against 46 TF/s for
// v_mfma_f64_16x16x4 (register-only issue loops).  Lane
// t = 16k + 4b + a supplies A_b(a, k) and lane t = 16k + 4b + c supplies
// B_b(k, c); lane 16a + 4b + c receives D_b(a, c).  Taking A_b = rows
// 4b..4b+3 of a 16-row strip and the same four columns as B_b for every
// block, one instruction produces a 16 x 4 strip of C from the fragments
// A(lane % 16, lane / 16) and B(lane / 16, 4q + lane % 4).  The products are
// the same four-term FP64 dot products as with the 16x16x4 form, and the
// results are bitwise identical to it.
// Two tilings: 64 x 128 per workgroup (each of 8 waves 32 x 32, 512 threads)
// for K >= 128 and N >= 128, 64 x 64 (each of 4 waves 32 x 32) otherwise.  FP32 keeps
// v_mfma_f32_16x16x4 on the 64 x 64 tiling.  The epilogue stages the tile
// through LDS so C is read and written along columns (coalesced): the rank-16
// updates of the blocked inverse are pure C streaming.
//
// Workgroup -> (item, tile) mapping is XCD-aware: hardware dispatches
// consecutive workgroup ids round-robin over the 8 XCDs, so the linear id is
// remapped such that all tiles of one item land on the same XCD and share its
// L2 (the A and B panels of an item are re-read by every tile).
#include <algorithm>
#include <atomic>
#include <cstdlib>
#include <mutex>
#include <type_traits>
#include <unordered_map>

#include "hn_gemm.h"
#include "hn_prof.h"

namespace hn {

namespace {

constexpr int BM = 64, BN = 64, BK = 16, NT = 256;
constexpr int PAD = 2;  // row padding (elements) of the LDS planes

template <typename R> struct V4;
template <> struct V4<double> { typedef double __attribute__((ext_vector_type(4))) t; };
template <> struct V4<float> { typedef float __attribute__((ext_vector_type(4))) t; };

__device__ inline V4<double>::t mfma(double a, double b, V4<double>::t c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ inline V4<float>::t mfma(float a, float b, V4<float>::t c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

template <typename R> struct CplxT;
template <> struct CplxT<double> { typedef double2 t; };
template <> struct CplxT<float> { typedef float2 t; };

// XCD-aware remap of a linear workgroup id in [0, total_pad) (total_pad a
// multiple of 8) to a logical id: logical ids [x*per, (x+1)*per) run on XCD x.
__device__ inline int xcd_remap(int wg, int total_pad) {
  int per = total_pad >> 3;
  return (wg & 7) * per + (wg >> 3);
}

// row of element r (0..3) of lane `lane` inside a 16 x 16 MFMA output tile:
// v_mfma_f64_16x16x4_f64 keeps D[row][col] in lane col + 16*(row%4), element
// row/4; v_mfma_f32_16x16x4_f32 in lane col + 16*(row/4), element row%4.
template <typename R> __device__ inline int out_row(int lane, int r) {
  return sizeof(R) == 8 ? (lane >> 4) + 4 * r : (lane >> 4) * 4 + r;
}

template <typename R> struct Smem {
  R Ar[BK][BM + PAD];
  R Ai[BK][BM + PAD];
  R Br[BK][BN + PAD];
  R Bi[BK][BN + PAD];
};

// USE only separates the uses (masked, sweep, inverse update) in profiles.
template <typename R, int USE>
__global__ __launch_bounds__(NT) void k_gemm(GemmArgs<R> p, int tiles_m, int tiles_n,
                                             int total, int total_pad) {
  typedef typename CplxT<R>::t CT;
  typedef typename V4<R>::t v4;
  const int logical = xcd_remap(blockIdx.x, total_pad);
  if (logical >= total) return;
  const int tiles = tiles_m * tiles_n;
  const int b = logical / tiles;
  const int t = logical - b * tiles;
  if (p.active && p.active[b] == 0) return;
  const int tm = t % tiles_m, tn = t / tiles_m;
  const int m0 = tm * BM, n0 = tn * BN;
  if ((m0 >= p.skip_r0 && m0 + BM <= p.skip_r1) || (n0 >= p.skip_c0 && n0 + BN <= p.skip_c1))
    return;  // wholly inside a skipped range

  __shared__ Smem<R> sm1;

  const CT *A = p.A + (long long)b * p.sA;
  const CT *B = p.B + (long long)b * p.sB;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  const int M = p.m, N = p.n, K = p.k;
  const bool a_n = p.opA == OP_N, b_n = p.opB == OP_N;
  const R sa = (p.opA == OP_C) ? R(-1) : R(1), sb = (p.opB == OP_C) ? R(-1) : R(1);

  // global -> register staging of one K step (4 complex of A and of B per thread)
  CT ra[4], rb[4];
  auto load = [&](int k0) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      int i, kk;
      if (a_n) { i = tid & 63; kk = (tid >> 6) + 4 * r; }
      else { kk = tid & 15; i = (tid >> 4) + 16 * r; }
      const int gi = m0 + i, gk = k0 + kk;
      CT v = {0, 0};
      if (gi < M && gk < K) v = a_n ? A[gi + (long long)gk * p.lda] : A[gk + (long long)gi * p.lda];
      ra[r] = v;
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      int j, kk;
      if (b_n) { kk = tid & 15; j = (tid >> 4) + 16 * r; }
      else { j = tid & 63; kk = (tid >> 6) + 4 * r; }
      const int gj = n0 + j, gk = k0 + kk;
      CT v = {0, 0};
      if (gj < N && gk < K) v = b_n ? B[gk + (long long)gj * p.ldb] : B[gj + (long long)gk * p.ldb];
      rb[r] = v;
    }
  };
  auto store = [&](Smem<R> &sm) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      int i, kk;
      if (a_n) { i = tid & 63; kk = (tid >> 6) + 4 * r; }
      else { kk = tid & 15; i = (tid >> 4) + 16 * r; }
      sm.Ar[kk][i] = ra[r].x;
      sm.Ai[kk][i] = sa * ra[r].y;
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      int j, kk;
      if (b_n) { kk = tid & 15; j = (tid >> 4) + 16 * r; }
      else { j = tid & 63; kk = (tid >> 6) + 4 * r; }
      sm.Br[kk][j] = rb[r].x;
      sm.Bi[kk][j] = sb * rb[r].y;
    }
  };

  // FP64: each wave's 32 x 32 = 2 row strips x 8 column quads of 4x4x4
  // instructions; FP32: 2 x 2 tiles of 16x16x4.
  constexpr bool M4 = sizeof(R) == 8;
  constexpr int NP = 2, NQ = M4 ? 8 : 2;
  typedef typename std::conditional<M4, R, v4>::type acc_t;
  acc_t cr[NP][NQ], ci[NP][NQ];
#pragma unroll
  for (int i = 0; i < NP; ++i)
#pragma unroll
    for (int j = 0; j < NQ; ++j) {
      cr[i][j] = acc_t{0};
      ci[i][j] = acc_t{0};
    }

  // C values of this thread's epilogue elements, fetched at the start of the
  // epilogue
  CT *C = p.C + (long long)b * p.sC;
  const bool useb = (p.beta.x != 0.0 || p.beta.y != 0.0);
  constexpr int EQ = (64 * 32) / NT;
  CT cpre[2][EQ];
  auto fetch_c = [&](int j) {
#pragma unroll
    for (int q = 0; q < EQ; ++q) {
      const int e = tid + q * NT;
      const int gr = m0 + (e & 63), sc = e >> 6;
      const int gc = n0 + (sc >> 4) * 32 + j * 16 + (sc & 15);
      cpre[j][q] = (useb && gr < M && gc < N) ? C[gr + (long long)gc * p.ldc] : CT{0, 0};
    }
  };
  if (K > 0) load(0);
  for (int k0 = 0; k0 < K; k0 += BK) {
    Smem<R> &sm = sm1;
    store(sm);
    __syncthreads();
    if (k0 + BK < K) load(k0 + BK);
#pragma unroll
    for (int ks = 0; ks < BK; ks += 4) {
      const int kr = ks + (lane >> 4), c16 = lane & 15;
      if constexpr (M4) {
        R ar[2], ai[2], nai[2], br[8], bi[8];
#pragma unroll
        for (int i = 0; i < 2; ++i) {
          ar[i] = sm.Ar[kr][wm * 32 + i * 16 + c16];
          ai[i] = sm.Ai[kr][wm * 32 + i * 16 + c16];
          nai[i] = -ai[i];
        }
#pragma unroll
        for (int q = 0; q < 8; ++q) {
          br[q] = sm.Br[kr][wn * 32 + q * 4 + (lane & 3)];
          bi[q] = sm.Bi[kr][wn * 32 + q * 4 + (lane & 3)];
        }
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int q = 0; q < 8; ++q) {
            cr[i][q] = __builtin_amdgcn_mfma_f64_4x4x4f64(ar[i], br[q], cr[i][q], 0, 0, 0);
            ci[i][q] = __builtin_amdgcn_mfma_f64_4x4x4f64(ar[i], bi[q], ci[i][q], 0, 0, 0);
          }
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int q = 0; q < 8; ++q) {
            cr[i][q] = __builtin_amdgcn_mfma_f64_4x4x4f64(nai[i], bi[q], cr[i][q], 0, 0, 0);
            ci[i][q] = __builtin_amdgcn_mfma_f64_4x4x4f64(ai[i], br[q], ci[i][q], 0, 0, 0);
          }
      } else {
        R ar[2], ai[2], nai[2], br[2], bi[2];
#pragma unroll
        for (int i = 0; i < 2; ++i) {
          ar[i] = sm.Ar[kr][wm * 32 + i * 16 + c16];
          ai[i] = sm.Ai[kr][wm * 32 + i * 16 + c16];
          nai[i] = -ai[i];
        }
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          br[j] = sm.Br[kr][wn * 32 + j * 16 + c16];
          bi[j] = sm.Bi[kr][wn * 32 + j * 16 + c16];
        }
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int j = 0; j < 2; ++j) {
            cr[i][j] = mfma(ar[i], br[j], cr[i][j]);
            ci[i][j] = mfma(ar[i], bi[j], ci[i][j]);
          }
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int j = 0; j < 2; ++j) {
            cr[i][j] = mfma(nai[i], bi[j], cr[i][j]);
            ci[i][j] = mfma(ai[i], br[j], ci[i][j]);
          }
      }
    }
    __syncthreads();
  }
  Smem<R> &sm = sm1;

  // ---- epilogue through LDS: the four waves stage a 64 x 32 slab (tile
  // columns j*16..j*16+15 of both wave columns) per pass, then every thread
  // reads/writes C along columns (coalesced) ----
  constexpr int EL = BM + 1;  // leading dimension of the staged slab
  static_assert(32 * EL <= 2 * BK * (BM + PAD), "epilogue slab does not fit");  // Ar+Ai, Br+Bi
  R *Er = &sm.Ar[0][0];
  R *Ei = &sm.Br[0][0];
  CT *C2 = (p.C2 && (!p.active2 || p.active2[b] != 0)) ? p.C2 + (long long)b * p.sC2 : nullptr;
  const double ar_ = p.alpha.x, ai_ = p.alpha.y;
  const bool useb2 = (p.beta2.x != 0.0 || p.beta2.y != 0.0);
  fetch_c(0);
  fetch_c(1);
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    if constexpr (M4) {
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int row = wm * 32 + i * 16 + 4 * ((lane >> 2) & 3) + (lane >> 4);
          const int sc = wn * 16 + 4 * q + (lane & 3);
          Er[sc * EL + row] = cr[i][4 * j + q];
          Ei[sc * EL + row] = ci[i][4 * j + q];
        }
    } else {
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = wm * 32 + i * 16 + out_row<R>(lane, r);
          const int sc = wn * 16 + (lane & 15);
          Er[sc * EL + row] = cr[i][j][r];
          Ei[sc * EL + row] = ci[i][j][r];
        }
    }
    __syncthreads();
    // slab column sc holds tile column (sc / 16) * 32 + j * 16 + sc % 16
#pragma unroll
    for (int q = 0; q < EQ; ++q) {
      const int e = tid + q * NT;
      const int row = e & 63, sc = e >> 6;
      const int gr = m0 + row;
      const int gc = n0 + (sc >> 4) * 32 + j * 16 + (sc & 15);
      if (gr < M && gc < N && !(gc >= p.skip_c0 && gc < p.skip_c1) &&
          !(gr >= p.skip_r0 && gr < p.skip_r1)) {
        const double xr = (double)Er[sc * EL + row], xi = (double)Ei[sc * EL + row];
        const double vr = ar_ * xr - ai_ * xi, vi = ar_ * xi + ai_ * xr;
        const long long o = gr + (long long)gc * p.ldc;
        double yr = vr, yi = vi;
        if (useb) {
          const CT c = cpre[j][q];
          yr += p.beta.x * (double)c.x - p.beta.y * (double)c.y;
          yi += p.beta.x * (double)c.y + p.beta.y * (double)c.x;
        }
        C[o] = CT{(R)yr, (R)yi};
        if (C2) {
          const long long o2 = gr + (long long)gc * p.ldc2;
          double zr = vr, zi = vi;
          if (useb2) {
            const CT c = C2[o2];
            zr += p.beta2.x * (double)c.x - p.beta2.y * (double)c.y;
            zi += p.beta2.x * (double)c.y + p.beta2.y * (double)c.x;
          }
          C2[o2] = CT{(R)zr, (R)zi};
        }
      }
    }
    __syncthreads();
  }
}

// ---- FP64 wide tile: 64 x 128 per workgroup of 512 threads, each of the
// eight waves 32 x 32 (2 row strips x 8 column quads of 4x4x4 instructions):
// two waves per SIMD, so one wave's LDS reads and barrier waits overlap the
// other's matrix-core work ----------------------------------------------------
constexpr int WBN = 128, WPA = 2, WPB = 2, NTW = 512;
struct SmemW {
  double Ar[BK][BM + WPA];
  double Ai[BK][BM + WPA];
  double Br[BK][WBN + WPB];
  double Bi[BK][WBN + WPB];
};
struct SmemW2 {
  SmemW buf[2];
};

template <int USE>
__global__ __launch_bounds__(NTW) void k_gemm_w(GemmArgs<double> p, int tiles_m, int tiles_n,
                                               int total, int total_pad) {
  typedef double2 CT;
  const int logical = xcd_remap(blockIdx.x, total_pad);
  if (logical >= total) return;
  const int tiles = tiles_m * tiles_n;
  const int b = logical / tiles;
  const int t = logical - b * tiles;
  if (p.active && p.active[b] == 0) return;
  const int tm = t % tiles_m, tn = t / tiles_m;
  const int m0 = tm * BM, n0 = tn * WBN;
  if ((m0 >= p.skip_r0 && m0 + BM <= p.skip_r1) || (n0 >= p.skip_c0 && n0 + WBN <= p.skip_c1))
    return;  // wholly inside a skipped range

  __shared__ SmemW2 sm2;

  const CT *A = p.A + (long long)b * p.sA;
  const CT *B = p.B + (long long)b * p.sB;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int M = p.m, N = p.n, K = p.k;
  const bool a_n = p.opA == OP_N, b_n = p.opB == OP_N;
  const double sa = (p.opA == OP_C) ? -1.0 : 1.0, sb = (p.opB == OP_C) ? -1.0 : 1.0;

  CT ra[2], rb[4];
  const int wm = wave >> 2, wn = wave & 3;
  const bool full_mn = (m0 + BM <= M) && (n0 + WBN <= N);
  auto load = [&](int k0) {
    if (full_mn && k0 + BK <= K) {  // interior tile: no per-element bounds tests
#pragma unroll
      for (int r = 0; r < 2; ++r) {
        int i, kk;
        if (a_n) { i = tid & 63; kk = (tid >> 6) + 8 * r; }
        else { kk = tid & 15; i = (tid >> 4) + 32 * r; }
        const int gi = m0 + i, gk = k0 + kk;
        ra[r] = a_n ? A[gi + (long long)gk * p.lda] : A[gk + (long long)gi * p.lda];
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        int j, kk;
        if (b_n) { kk = tid & 15; j = (tid >> 4) + 32 * r; }
        else { j = tid & 127; kk = (tid >> 7) + 4 * r; }
        const int gj = n0 + j, gk = k0 + kk;
        rb[r] = b_n ? B[gk + (long long)gj * p.ldb] : B[gj + (long long)gk * p.ldb];
      }
      return;
    }
#pragma unroll
    for (int r = 0; r < 2; ++r) {
      int i, kk;
      if (a_n) { i = tid & 63; kk = (tid >> 6) + 8 * r; }
      else { kk = tid & 15; i = (tid >> 4) + 32 * r; }
      const int gi = m0 + i, gk = k0 + kk;
      CT v = {0, 0};
      if (gi < M && gk < K) v = a_n ? A[gi + (long long)gk * p.lda] : A[gk + (long long)gi * p.lda];
      ra[r] = v;
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      int j, kk;
      if (b_n) { kk = tid & 15; j = (tid >> 4) + 32 * r; }
      else { j = tid & 127; kk = (tid >> 7) + 4 * r; }
      const int gj = n0 + j, gk = k0 + kk;
      CT v = {0, 0};
      if (gj < N && gk < K) v = b_n ? B[gk + (long long)gj * p.ldb] : B[gj + (long long)gk * p.ldb];
      rb[r] = v;
    }
  };
  auto store = [&](SmemW &sm) {
#pragma unroll
    for (int r = 0; r < 2; ++r) {
      int i, kk;
      if (a_n) { i = tid & 63; kk = (tid >> 6) + 8 * r; }
      else { kk = tid & 15; i = (tid >> 4) + 32 * r; }
      sm.Ar[kk][i] = ra[r].x;
      sm.Ai[kk][i] = sa * ra[r].y;
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      int j, kk;
      if (b_n) { kk = tid & 15; j = (tid >> 4) + 32 * r; }
      else { j = tid & 127; kk = (tid >> 7) + 4 * r; }
      sm.Br[kk][j] = rb[r].x;
      sm.Bi[kk][j] = sb * rb[r].y;
    }
  };

  double cr[2][8], ci[2][8];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      cr[i][q] = 0.0;
      ci[i][q] = 0.0;
    }
  // C values of this thread's epilogue elements, fetched at the start of the
  // epilogue (a prefetch before the K loop would not fit the registers)
  CT *C = p.C + (long long)b * p.sC;
  const bool useb = (p.beta.x != 0.0 || p.beta.y != 0.0);
  constexpr int EQ = (64 * 64) / NTW;
  CT cpre[2][EQ];
  auto fetch_c = [&](int j) {
#pragma unroll
    for (int q = 0; q < EQ; ++q) {
      const int e = tid + q * NTW;
      const int gr = m0 + (e & 63), sc = e >> 6;
      const int gc = n0 + (sc >> 4) * 32 + j * 16 + (sc & 15);
      cpre[j][q] = (useb && gr < M && gc < N) ? C[gr + (long long)gc * p.ldc] : CT{0, 0};
    }
  };
  int cur = 0;
  if (K > 0) {
    load(0);
    store(sm2.buf[0]);
    __syncthreads();
    if (BK < K) load(BK);
  }
  for (int k0 = 0; k0 < K; k0 += BK) {
    SmemW &sm = sm2.buf[cur];
    // operand fragments of sub-step s + 1 are read from LDS while the MFMAs of
    // sub-step s run (two register sets)
    double fa[2][2], fai[2][2], fb[2][8], fbi[2][8];
    auto frag = [&](int ks, int slot) {
      const int kr = ks + (lane >> 4), c16 = lane & 15;
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        fa[slot][i] = sm.Ar[kr][wm * 32 + i * 16 + c16];
        fai[slot][i] = sm.Ai[kr][wm * 32 + i * 16 + c16];
      }
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        fb[slot][q] = sm.Br[kr][wn * 32 + q * 4 + (lane & 3)];
        fbi[slot][q] = sm.Bi[kr][wn * 32 + q * 4 + (lane & 3)];
      }
    };
    frag(0, 0);
#pragma unroll
    for (int s = 0; s < BK / 4; ++s) {
      const int sl = s & 1;
      if (s + 1 < BK / 4) frag(4 * (s + 1), sl ^ 1);
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int q = 0; q < 8; ++q) {
          cr[i][q] = __builtin_amdgcn_mfma_f64_4x4x4f64(fa[sl][i], fb[sl][q], cr[i][q], 0, 0, 0);
          ci[i][q] = __builtin_amdgcn_mfma_f64_4x4x4f64(fa[sl][i], fbi[sl][q], ci[i][q], 0, 0, 0);
        }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int q = 0; q < 8; ++q) {
          cr[i][q] = __builtin_amdgcn_mfma_f64_4x4x4f64(-fai[sl][i], fbi[sl][q], cr[i][q], 0, 0, 0);
          ci[i][q] = __builtin_amdgcn_mfma_f64_4x4x4f64(fai[sl][i], fb[sl][q], ci[i][q], 0, 0, 0);
        }
    }
    if (k0 + BK < K) {
      store(sm2.buf[cur ^ 1]);
      if (k0 + 2 * BK < K) load(k0 + 2 * BK);
    }
    __syncthreads();
    cur ^= 1;
  }

  // ---- epilogue: two passes of a 64 x 64 slab (columns 16j..16j+15 of every wave) ----
  constexpr int EL = BM + 1;
  static_assert(2 * 64 * EL * sizeof(double) <= sizeof(SmemW2), "epilogue slab does not fit");
  double *Er = &sm2.buf[0].Ar[0][0];
  double *Ei = Er + 64 * EL;
  CT *C2 = (p.C2 && (!p.active2 || p.active2[b] != 0)) ? p.C2 + (long long)b * p.sC2 : nullptr;
  const double ar_ = p.alpha.x, ai_ = p.alpha.y;
  const bool useb2 = (p.beta2.x != 0.0 || p.beta2.y != 0.0);
  fetch_c(0);
  fetch_c(1);
#pragma unroll
  for (int j = 0; j < 2; ++j) {
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int row = wm * 32 + i * 16 + 4 * ((lane >> 2) & 3) + (lane >> 4);
        const int sc = wn * 16 + 4 * q + (lane & 3);
        Er[sc * EL + row] = cr[i][4 * j + q];
        Ei[sc * EL + row] = ci[i][4 * j + q];
      }
    __syncthreads();
    // slab column sc holds tile column (sc / 16) * 32 + j * 16 + sc % 16
#pragma unroll
    for (int q = 0; q < EQ; ++q) {
      const int e = tid + q * NTW;
      const int row = e & 63, sc = e >> 6;
      const int gr = m0 + row;
      const int gc = n0 + (sc >> 4) * 32 + j * 16 + (sc & 15);
      if (gr < M && gc < N && !(gc >= p.skip_c0 && gc < p.skip_c1) &&
          !(gr >= p.skip_r0 && gr < p.skip_r1)) {
        const double xr = Er[sc * EL + row], xi = Ei[sc * EL + row];
        const double vr = ar_ * xr - ai_ * xi, vi = ar_ * xi + ai_ * xr;
        const long long o = gr + (long long)gc * p.ldc;
        double yr = vr, yi = vi;
        if (useb) {
          const CT c = cpre[j][q];
          yr += p.beta.x * c.x - p.beta.y * c.y;
          yi += p.beta.x * c.y + p.beta.y * c.x;
        }
        C[o] = CT{yr, yi};
        if (C2) {
          const long long o2 = gr + (long long)gc * p.ldc2;
          double zr = vr, zi = vi;
          if (useb2) {
            const CT c = C2[o2];
            zr += p.beta2.x * c.x - p.beta2.y * c.y;
            zi += p.beta2.x * c.y + p.beta2.y * c.x;
          }
          C2[o2] = CT{zr, zi};
        }
      }
    }
    __syncthreads();
  }
}

// ---- FP64 single-wave tile: 32 x 32 per 64-thread workgroup (the work of
// one wave of the 64 x 64 tile, same fragments and the same MFMA order, so
// results are bitwise identical to it).  Workgroups are independent (no
// barrier couples four waves), so at two waves per SIMD eight tiles per CU
// hide each other's operand and C latencies; the epilogue writes the
// accumulators straight to memory (16 rows x 4 columns per instruction).
constexpr int SBT = 32, NTS = 64;
struct SmemS {
  double Ar[BK][SBT + PAD];
  double Ai[BK][SBT + PAD];
  double Br[BK][SBT + PAD];
  double Bi[BK][SBT + PAD];
};

template <int USE>
__global__ __launch_bounds__(NTS) void k_gemm_s(GemmArgs<double> p, int tiles_m, int tiles_n,
                                               int total, int total_pad) {
  typedef double2 CT;
  const int logical = xcd_remap(blockIdx.x, total_pad);
  if (logical >= total) return;
  const int tiles = tiles_m * tiles_n;
  const int b = logical / tiles;
  const int t = logical - b * tiles;
  if (p.active && p.active[b] == 0) return;
  const int tm = t % tiles_m, tn = t / tiles_m;
  const int m0 = tm * SBT, n0 = tn * SBT;
  if ((m0 >= p.skip_r0 && m0 + SBT <= p.skip_r1) || (n0 >= p.skip_c0 && n0 + SBT <= p.skip_c1))
    return;  // wholly inside a skipped range

  __shared__ SmemS sm;
  const CT *A = p.A + (long long)b * p.sA;
  const CT *B = p.B + (long long)b * p.sB;
  const int lane = threadIdx.x;
  const int M = p.m, N = p.n, K = p.k;
  const bool a_n = p.opA == OP_N, b_n = p.opB == OP_N;
  const double sa = (p.opA == OP_C) ? -1.0 : 1.0, sb = (p.opB == OP_C) ? -1.0 : 1.0;

  // global -> register staging of one K step (8 complex of A and of B per lane)
  CT ra[8], rb[8];
  auto load = [&](int k0) {
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      int i, kk;
      if (a_n) { i = lane & 31; kk = (lane >> 5) + 2 * r; }
      else { kk = lane & 15; i = (lane >> 4) + 4 * r; }
      const int gi = m0 + i, gk = k0 + kk;
      CT v = {0, 0};
      if (gi < M && gk < K) v = a_n ? A[gi + (long long)gk * p.lda] : A[gk + (long long)gi * p.lda];
      ra[r] = v;
    }
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      int j, kk;
      if (b_n) { kk = lane & 15; j = (lane >> 4) + 4 * r; }
      else { j = lane & 31; kk = (lane >> 5) + 2 * r; }
      const int gj = n0 + j, gk = k0 + kk;
      CT v = {0, 0};
      if (gj < N && gk < K) v = b_n ? B[gk + (long long)gj * p.ldb] : B[gj + (long long)gk * p.ldb];
      rb[r] = v;
    }
  };
  auto store = [&]() {
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      int i, kk;
      if (a_n) { i = lane & 31; kk = (lane >> 5) + 2 * r; }
      else { kk = lane & 15; i = (lane >> 4) + 4 * r; }
      sm.Ar[kk][i] = ra[r].x;
      sm.Ai[kk][i] = sa * ra[r].y;
    }
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      int j, kk;
      if (b_n) { kk = lane & 15; j = (lane >> 4) + 4 * r; }
      else { j = lane & 31; kk = (lane >> 5) + 2 * r; }
      sm.Br[kk][j] = rb[r].x;
      sm.Bi[kk][j] = sb * rb[r].y;
    }
  };

  double cr[2][8], ci[2][8];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int q = 0; q < 8; ++q) { cr[i][q] = 0.0; ci[i][q] = 0.0; }

  if (K > 0) load(0);
  for (int k0 = 0; k0 < K; k0 += BK) {
    store();
    __syncthreads();
    if (k0 + BK < K) load(k0 + BK);
#pragma unroll
    for (int ks = 0; ks < BK; ks += 4) {
      const int kr = ks + (lane >> 4), c16 = lane & 15;
      double ar[2], ai[2], nai[2], br[8], bi[8];
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        ar[i] = sm.Ar[kr][i * 16 + c16];
        ai[i] = sm.Ai[kr][i * 16 + c16];
        nai[i] = -ai[i];
      }
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        br[q] = sm.Br[kr][q * 4 + (lane & 3)];
        bi[q] = sm.Bi[kr][q * 4 + (lane & 3)];
      }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int q = 0; q < 8; ++q) {
          cr[i][q] = __builtin_amdgcn_mfma_f64_4x4x4f64(ar[i], br[q], cr[i][q], 0, 0, 0);
          ci[i][q] = __builtin_amdgcn_mfma_f64_4x4x4f64(ar[i], bi[q], ci[i][q], 0, 0, 0);
        }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int q = 0; q < 8; ++q) {
          cr[i][q] = __builtin_amdgcn_mfma_f64_4x4x4f64(nai[i], bi[q], cr[i][q], 0, 0, 0);
          ci[i][q] = __builtin_amdgcn_mfma_f64_4x4x4f64(ai[i], br[q], ci[i][q], 0, 0, 0);
        }
    }
    __syncthreads();
  }

  // ---- epilogue: lane holds rows row0 + 16 i, columns col0 + 4 q ----
  CT *C = p.C + (long long)b * p.sC;
  CT *C2 = (p.C2 && (!p.active2 || p.active2[b] != 0)) ? p.C2 + (long long)b * p.sC2 : nullptr;
  const bool useb = (p.beta.x != 0.0 || p.beta.y != 0.0);
  const bool useb2 = (p.beta2.x != 0.0 || p.beta2.y != 0.0);
  const double ar_ = p.alpha.x, ai_ = p.alpha.y;
  const int row0 = m0 + 4 * ((lane >> 2) & 3) + (lane >> 4);
  const int col0 = n0 + (lane & 3);
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    CT cv[8];
    const int gr = row0 + i * 16;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int gc = col0 + 4 * q;
      cv[q] = (useb && gr < M && gc < N) ? C[gr + (long long)gc * p.ldc] : CT{0, 0};
    }
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int gc = col0 + 4 * q;
      if (gr < M && gc < N && !(gc >= p.skip_c0 && gc < p.skip_c1) &&
          !(gr >= p.skip_r0 && gr < p.skip_r1)) {
        const double xr = cr[i][q], xi = ci[i][q];
        const double vr = ar_ * xr - ai_ * xi, vi = ar_ * xi + ai_ * xr;
        double yr = vr, yi = vi;
        if (useb) {
          const CT c = cv[q];
          yr += p.beta.x * c.x - p.beta.y * c.y;
          yi += p.beta.x * c.y + p.beta.y * c.x;
        }
        C[gr + (long long)gc * p.ldc] = CT{yr, yi};
        if (C2) {
          const long long o2 = gr + (long long)gc * p.ldc2;
          double zr = vr, zi = vi;
          if (useb2) {
            const CT c = C2[o2];
            zr += p.beta2.x * c.x - p.beta2.y * c.y;
            zi += p.beta2.x * c.y + p.beta2.y * c.x;
          }
          C2[o2] = CT{zr, zi};
        }
      }
    }
  }
}

// V: 0 the 64 x 64 tile, 1 the 64 x 128 tile, 2 the 32 x 32 single-wave tile
// (FP64 only; FP32 always takes the 64 x 64 tile)
template <typename R>
void run(int V, int use, const GemmArgs<R> &a, hipStream_t s) {
  const bool W = V == 1, S1 = V == 2;
  const int tr = S1 ? SBT : BM, tc = W ? WBN : S1 ? SBT : BN;
  int tiles_m = (a.m + tr - 1) / tr, tiles_n = (a.n + tc - 1) / tc;
  long long total = (long long)tiles_m * tiles_n * a.batch;
  if (total > (1LL << 30)) throw std::runtime_error("gemm: grid too large");
  int total_pad = (int)((total + 7) / 8 * 8);
  if constexpr (sizeof(R) == 8) {
    if (W) {
      if (use == PT_GEMM_SWEEP)
        hipLaunchKernelGGL((k_gemm_w<PT_GEMM_SWEEP>), dim3(total_pad), dim3(NTW), 0, s, a, tiles_m,
                           tiles_n, (int)total, total_pad);
      else if (use == PT_GEMM_INV)
        hipLaunchKernelGGL((k_gemm_w<PT_GEMM_INV>), dim3(total_pad), dim3(NTW), 0, s, a, tiles_m,
                           tiles_n, (int)total, total_pad);
      else
        hipLaunchKernelGGL((k_gemm_w<PT_GEMM_MASKED>), dim3(total_pad), dim3(NTW), 0, s, a, tiles_m,
                           tiles_n, (int)total, total_pad);
      HN_CHECK(hipGetLastError());
      return;
    }
    if (S1) {
      if (use == PT_GEMM_SWEEP)
        hipLaunchKernelGGL((k_gemm_s<PT_GEMM_SWEEP>), dim3(total_pad), dim3(NTS), 0, s, a, tiles_m,
                           tiles_n, (int)total, total_pad);
      else if (use == PT_GEMM_INV)
        hipLaunchKernelGGL((k_gemm_s<PT_GEMM_INV>), dim3(total_pad), dim3(NTS), 0, s, a, tiles_m,
                           tiles_n, (int)total, total_pad);
      else
        hipLaunchKernelGGL((k_gemm_s<PT_GEMM_MASKED>), dim3(total_pad), dim3(NTS), 0, s, a, tiles_m,
                           tiles_n, (int)total, total_pad);
      HN_CHECK(hipGetLastError());
      return;
    }
  }
  if (use == PT_GEMM_SWEEP)
    hipLaunchKernelGGL((k_gemm<R, PT_GEMM_SWEEP>), dim3(total_pad), dim3(NT), 0, s, a, tiles_m,
                       tiles_n, (int)total, total_pad);
  else if (use == PT_GEMM_INV)
    hipLaunchKernelGGL((k_gemm<R, PT_GEMM_INV>), dim3(total_pad), dim3(NT), 0, s, a, tiles_m,
                       tiles_n, (int)total, total_pad);
  else
    hipLaunchKernelGGL((k_gemm<R, PT_GEMM_MASKED>), dim3(total_pad), dim3(NT), 0, s, a, tiles_m,
                       tiles_n, (int)total, total_pad);
  HN_CHECK(hipGetLastError());
}

// ---- tile-shape autotuner (FP64) ------------------------------------------
// The first launch of each shape class -- (m, n, k, operand forms, beta and
// dual-output and mask presence, skipped ranges, use, batch rounded down to
// a power of two) -- times the 64 x 64 four-wave tile, the 64 x 128
// eight-wave tile (n >= 64) and the 32 x 32 single-wave tile on the real
// operands, writing into a scratch output, and keeps the fixed rule's tile
// unless another is at least 3 % faster, for every later launch of that
// class.  The tilings run the same MFMA chains in the same order, so the
// choice never changes a result.  Off (the fixed rule: wide for k >= 64 and n >= 128) while a
// stream is being captured into a graph, with HIPNEGF_GEMM_TUNE=0, or when a
// variant is forced (gemm_force_variant).
struct TuneKey {
  int m, n, k, opA, opB, flags, use, bb;
  bool operator==(const TuneKey &o) const {
    return m == o.m && n == o.n && k == o.k && opA == o.opA && opB == o.opB && flags == o.flags &&
           use == o.use && bb == o.bb;
  }
};
struct TuneKeyHash {
  size_t operator()(const TuneKey &k) const {
    size_t h = (size_t)k.m * 1000003u ^ (size_t)k.n * 10007u ^ (size_t)k.k * 101u;
    return h ^ ((size_t)k.opA << 40) ^ ((size_t)k.opB << 44) ^ ((size_t)k.flags << 48) ^
           ((size_t)k.use << 56) ^ ((size_t)k.bb << 60);
  }
};
std::mutex g_tune_mu;
std::unordered_map<TuneKey, int, TuneKeyHash> g_tune;  // 0: 64 x 64, 1: 64 x 128
std::atomic<int> g_force{-1};

bool tune_on() {
  static const int on = [] {
    const char *e = std::getenv("HIPNEGF_GEMM_TUNE");
    return e ? std::atoi(e) : 1;
  }();
  return on != 0;
}

int pick_variant(const GemmArgs<double> &a, int use, hipStream_t s) {
  const int f = g_force.load();
  if (f >= 0) return f;
  const int rule = a.k >= 64 && a.n >= 128 ? 1 : 0;
  if (a.n < 16 || a.m < 16 || a.k < 16 || !tune_on()) return rule;
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(s, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) return rule;
  int bb = 0;
  while ((2 << bb) <= a.batch) ++bb;
  const int flags = ((a.beta.x != 0.0 || a.beta.y != 0.0) ? 1 : 0) | (a.C2 ? 2 : 0) |
                    (a.active ? 4 : 0) | (a.active2 ? 8 : 0) |
                    ((a.skip_c1 > a.skip_c0 || a.skip_r1 > a.skip_r0) ? 16 : 0) |
                    (a.sA == 0 ? 32 : 0) | (a.sB == 0 ? 64 : 0);
  const TuneKey key{a.m, a.n, a.k, a.opA, a.opB, flags, use, bb};
  {
    std::lock_guard<std::mutex> g(g_tune_mu);
    auto it = g_tune.find(key);
    if (it != g_tune.end()) return it->second;
  }
  // scratch outputs of the same layout (the real C and C2 are not touched)
  auto span = [&](long long sC, int ldc) {
    return (long long)(a.batch - 1) * sC + (long long)ldc * (a.n - 1) + a.m;
  };
  GemmArgs<double> t = a;
  double2 *c1 = nullptr, *c2 = nullptr;
  // without room for the scratch outputs (a caller near its HBM budget, or
  // less than twice their size plus 1 GiB free) this launch takes the fixed
  // rule untuned; a later one may tune the class
  const size_t need = sizeof(double2) * ((size_t)span(a.sC, a.ldc) + (a.C2 ? (size_t)span(a.sC2, a.ldc2) : 0));
  size_t free_b = 0, total_b = 0;
  if (hipMemGetInfo(&free_b, &total_b) != hipSuccess || 2 * need + (1ull << 30) > free_b) {
    (void)hipGetLastError();
    return rule;
  }
  if (hipMallocAsync(&c1, sizeof(double2) * span(a.sC, a.ldc), s) != hipSuccess) {
    (void)hipGetLastError();
    return rule;
  }
  t.C = c1;
  if (a.C2) {
    if (hipMallocAsync(&c2, sizeof(double2) * span(a.sC2, a.ldc2), s) != hipSuccess) {
      (void)hipGetLastError();
      HN_CHECK(hipFreeAsync(c1, s));
      return rule;
    }
    t.C2 = c2;
  }
  hipEvent_t e0, e1;
  HN_CHECK(hipEventCreate(&e0));
  HN_CHECK(hipEventCreate(&e1));
  float ms[3] = {0.0f, 0.0f, 0.0f};
  for (int v = 0; v < 3; ++v) {
    if (v == 1 && a.n < 64) {  // the 64 x 128 tile would be mostly empty
      ms[v] = 1e30f;
      continue;
    }
    run<double>(v, use, t, s);  // first use of the code object
    HN_CHECK(hipEventRecord(e0, s));
    run<double>(v, use, t, s);
    run<double>(v, use, t, s);
    HN_CHECK(hipEventRecord(e1, s));
    HN_CHECK(hipEventSynchronize(e1));
    HN_CHECK(hipEventElapsedTime(&ms[v], e0, e1));
  }
  HN_CHECK(hipEventDestroy(e0));
  HN_CHECK(hipEventDestroy(e1));
  HN_CHECK(hipFreeAsync(c1, s));
  if (c2) HN_CHECK(hipFreeAsync(c2, s));
  // the fixed rule's tile unless another is clearly faster (3 %): shapes
  // where the tiles are within timing noise keep one choice from run to run
  int best = 0;
  for (int v = 1; v < 3; ++v)
    if (ms[v] < ms[best]) best = v;
  const int choice = ms[best] < 0.97f * ms[rule] ? best : rule;
  std::lock_guard<std::mutex> g(g_tune_mu);
  g_tune[key] = choice;
  return choice;
}

template <typename R>
void launch(const GemmArgs<R> &a, hipStream_t s) {
  if (a.m <= 0 || a.n <= 0 || a.batch <= 0) return;
  const int use = a.tag >= 0 ? a.tag : (a.active ? PT_GEMM_MASKED : PT_GEMM_SWEEP);
  int V = 0;
  if constexpr (sizeof(R) == 8) V = pick_variant(a, use, s);
  // algorithmic work: 8 real flops per complex multiply-add over the columns written
  const int skipped = std::max(0, std::min(a.skip_c1, a.n) - std::max(a.skip_c0, 0));
  const int skipped_r = std::max(0, std::min(a.skip_r1, a.m) - std::max(a.skip_r0, 0));
  const double ncol = (double)(a.n - skipped) * (double)(a.m - skipped_r) / (double)a.m;
  const double items = a.prof_items >= 0 ? (double)a.prof_items
                      : a.prof_items == -2 ? 1.0
                      : ((a.active || a.active2) ? -1.0 : (double)a.batch);
  const double flops = items < 0 ? -1.0 : 8.0 * a.m * ncol * a.k * items;
  const double cb = (double)sizeof(typename GemmArgs<R>::CT);
  const bool rc = a.beta.x != 0.0 || a.beta.y != 0.0;
  const double nb = items < 0 ? (double)a.batch : items;
  const double bytes = cb * ((a.sA ? nb : 1.0) * a.m * a.k + (a.sB ? nb : 1.0) * a.k * a.n +
                             (rc ? 2.0 : 1.0) * nb * a.m * ncol);
  // profile variant: 0 the 64 x 64 FP64 tile, 1 64 x 128, 2 FP32, 3 the 32 x 32 single-wave tile
  ProfScope ps(use, flops, bytes, s, sizeof(R) == 8 ? (V == 2 ? 3 : V) : 2);
  if (a.prof_token) *a.prof_token = ps.tok;
  run<R>(V, use, a, s);
}

}  // namespace

void gemm_z(const GemmArgs<double> &a, hipStream_t s) { launch<double>(a, s); }
void gemm_c(const GemmArgs<float> &a, hipStream_t s) { launch<float>(a, s); }

void gemm_force_variant(int v) { g_force.store(v < 0 || v > 2 ? -1 : v); }

void gemm_tune_stats(long long out[4]) {
  std::lock_guard<std::mutex> g(g_tune_mu);
  out[0] = (long long)g_tune.size();
  out[1] = out[3] = 0;
  for (auto &kv : g_tune) {
    out[1] += kv.second == 1;
    out[3] += kv.second == 2;
  }
  out[2] = tune_on() ? 1 : 0;
}

}  // namespace hn
