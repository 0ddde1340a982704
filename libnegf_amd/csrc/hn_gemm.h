// hn_gemm.h -- batched complex GEMM on CDNA4 matrix cores (host interface).
#pragma once
#include "hn_common.h"

namespace hn {

// C_b = alpha op(A_b) op(B_b) + beta C_b     for b in [0, batch)
// (optional) C2_b = alpha op(A_b) op(B_b) + beta2 C2_b  (dual output: the
// decimation subtracts one product from two matrices, contselfenergy.F90:312-314)
//
// Operands of item b live at X + b * strideX (strideX may be 0: shared operand).
// Items with active[b] == 0 are skipped (per-item convergence masks).
// Columns [skip_c0, skip_c1) and rows [skip_r0, skip_r1) of C are left
// untouched (the super-panel of the blocked inverse); tiles that lie wholly
// inside either range are not computed.
template <typename R>
struct GemmArgs {
  typedef typename std::conditional<sizeof(R) == 8, double2, float2>::type CT;
  int m = 0, n = 0, k = 0;
  const CT *A = nullptr; long long sA = 0; int lda = 0; int opA = OP_N;
  const CT *B = nullptr; long long sB = 0; int ldb = 0; int opB = OP_N;
  CT *C = nullptr; long long sC = 0; int ldc = 0;
  CT *C2 = nullptr; long long sC2 = 0; int ldc2 = 0;
  double2 alpha = {1.0, 0.0}, beta = {0.0, 0.0}, beta2 = {0.0, 0.0};
  const int *active = nullptr;
  const int *active2 = nullptr;  // optional separate mask of the second output C2
  int skip_c0 = 0, skip_c1 = 0;
  int skip_r0 = 0, skip_r1 = 0;
  int batch = 1;
  int tag = -1;  // profile tag (ProfTag); -1: masked if `active` is set, else sweep
  int prof_items = -1;  // items known to be active (for flop counts of masked launches);
                        // -2: count later (flops recorded per item, see Profiler::scale)
  int *prof_token = nullptr;  // out: profile record of this launch (-1 if none)
};

void gemm_z(const GemmArgs<double> &a, hipStream_t s);  // complex FP64
void gemm_c(const GemmArgs<float> &a, hipStream_t s);   // complex FP32 (FP32 accumulation)
// FP64 tile choice: -1 autotuned (default), 0 always the 64 x 64 tile, 1 always 64 x 128,
// 2 always the 32 x 32 single-wave tile
void gemm_force_variant(int v);
// {shape classes tuned, of which took the 64 x 128 tile, tuner on, of which took 32 x 32}
void gemm_tune_stats(long long out[4]);

}  // namespace hn
