// hn_prof.h -- optional per-launch timing of the engine's kernels.
//
// When enabled, a tagged launch site records a HIP event pair on the stream it
// launches on, with the launch's algorithmic flop count (negative: not
// counted, e.g. launches that skip converged items) and algorithmic bytes.
// Summaries per tag give the average launch duration and the achieved rate.
#pragma once
#include <mutex>
#include <vector>

#include "hn_common.h"

namespace hn {

enum ProfTag : int {
  PT_GEMM_MASKED = 0,  // batched GEMMs with per-item activity masks (decimation, SCBA)
  PT_GEMM_SWEEP = 1,   // batched GEMMs of the RGF sweeps and traces
  PT_GEMM_INV = 2,     // trailing updates of the blocked inverse
  PT_INV_PANEL = 3,    // panel factorisation of the blocked inverse
  PT_INV_SMALL = 4,    // register- or LDS-resident small-block inverse
  PT_FUSED = 5,        // whole small-block transmission solves in one launch
  PT_COUNT = 6
};

class Profiler {
 public:
  void enable(bool on);
  bool on() const { return on_; }
  // variant: which kernel served the tag (GEMMs: 0 the 64 x 64 FP64 tile,
  // 1 the 64 x 128 FP64 tile, 2 the FP32 tile; 0 elsewhere)
  int begin(int tag, double flops, double bytes, hipStream_t s, int variant = 0);
  void end(int token, hipStream_t s);
  // multiply a launch's recorded flops and bytes (recorded per item) by its item count
  void scale(int token, double items);
  // Waits for the recorded events; sums over launches of `tag` (and of
  // `variant`, unless it is negative).
  void summary(int tag, double *ms, double *flops, double *bytes, long long *launches,
               int variant = -1);
  void clear();

 private:
  struct Rec { int tag, variant; double flops, bytes; hipEvent_t a, b; };
  bool on_ = false;
  std::mutex mu_;
  std::vector<Rec> recs_;
  std::vector<hipEvent_t> pool_;
  hipEvent_t get();
};

Profiler &profiler();

// RAII helper for one launch
struct ProfScope {
  int tok = -1;
  hipStream_t s;
  ProfScope(int tag, double flops, double bytes, hipStream_t st, int variant = 0) : s(st) {
    if (profiler().on()) tok = profiler().begin(tag, flops, bytes, st, variant);
  }
  ~ProfScope() {
    if (tok >= 0) profiler().end(tok, s);
  }
};

// FP64 matrix-core issue loop (register operands only) on the current device:
// form 0 v_mfma_f64_4x4x4, 1 v_mfma_f64_16x16x4; returns TFLOP/s.
double mfma_issue_rate(int form, int iters, int blocks);

}  // namespace hn
