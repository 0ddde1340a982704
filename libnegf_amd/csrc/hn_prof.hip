// hn_prof.hip -- launch timing with HIP events (see hn_prof.h).
#include "hn_prof.h"

namespace hn {

Profiler &profiler() {
  static Profiler p;
  return p;
}

void Profiler::enable(bool on) {
  std::lock_guard<std::mutex> g(mu_);
  on_ = on;
}

hipEvent_t Profiler::get() {
  if (!pool_.empty()) {
    hipEvent_t e = pool_.back();
    pool_.pop_back();
    return e;
  }
  hipEvent_t e;
  HN_CHECK(hipEventCreate(&e));
  return e;
}

int Profiler::begin(int tag, double flops, double bytes, hipStream_t s, int variant) {
  std::lock_guard<std::mutex> g(mu_);
  if (!on_) return -1;
  Rec r{tag, variant, flops, bytes, get(), get()};
  HN_CHECK(hipEventRecord(r.a, s));
  recs_.push_back(r);
  return (int)recs_.size() - 1;
}

void Profiler::end(int token, hipStream_t s) {
  std::lock_guard<std::mutex> g(mu_);
  if (token < 0 || token >= (int)recs_.size()) return;
  (void)hipEventRecord(recs_[token].b, s);
}

void Profiler::scale(int token, double items) {
  std::lock_guard<std::mutex> g(mu_);
  if (token < 0 || token >= (int)recs_.size()) return;
  recs_[token].flops *= items;
  recs_[token].bytes *= items;
}

void Profiler::summary(int tag, double *ms, double *flops, double *bytes, long long *launches,
                       int variant) {
  std::lock_guard<std::mutex> g(mu_);
  double t = 0.0, f = 0.0, by = 0.0;
  long long n = 0;
  bool f_known = true;
  for (const Rec &r : recs_) {
    if (r.tag != tag || (variant >= 0 && r.variant != variant)) continue;
    HN_CHECK(hipEventSynchronize(r.b));
    float e = 0.0f;
    HN_CHECK(hipEventElapsedTime(&e, r.a, r.b));
    t += e;
    if (r.flops < 0) f_known = false;
    f += r.flops;
    by += r.bytes;
    ++n;
  }
  if (ms) *ms = t;
  if (flops) *flops = f_known ? f : -1.0;
  if (bytes) *bytes = by;
  if (launches) *launches = n;
}

void Profiler::clear() {
  std::lock_guard<std::mutex> g(mu_);
  for (const Rec &r : recs_) {
    (void)hipEventSynchronize(r.b);
    pool_.push_back(r.a);
    pool_.push_back(r.b);
  }
  recs_.clear();
}

// ---------------------------------------------------------------------------
// Matrix-core issue loop (counter calibration): every wave runs NC independent
// accumulation chains of one FP64 MFMA form with register operands only, so
// the instruction rate is the form's peak on this GPU and a hardware counter
// read over it has a known utilisation.
// ---------------------------------------------------------------------------
namespace {
typedef double v4d __attribute__((ext_vector_type(4)));
constexpr int kCalibChains = 8;

__global__ __launch_bounds__(256) void k_mfma_issue_4x4x4(int iters, double *out) {
  const double a = 1.0 + 1e-9 * threadIdx.x, b = 1.0 - 1e-9 * blockIdx.x;
  double acc[kCalibChains];
#pragma unroll
  for (int c = 0; c < kCalibChains; ++c) acc[c] = 0.0;
  for (int i = 0; i < iters; ++i)
#pragma unroll
    for (int c = 0; c < kCalibChains; ++c) acc[c] = __builtin_amdgcn_mfma_f64_4x4x4f64(a, b, acc[c], 0, 0, 0);
  double s = 0.0;
#pragma unroll
  for (int c = 0; c < kCalibChains; ++c) s += acc[c];
  out[blockIdx.x * 256 + threadIdx.x] = s;
}

__global__ __launch_bounds__(256) void k_mfma_issue_16x16x4(int iters, double *out) {
  const double a = 1.0 + 1e-9 * threadIdx.x, b = 1.0 - 1e-9 * blockIdx.x;
  v4d acc[kCalibChains];
#pragma unroll
  for (int c = 0; c < kCalibChains; ++c) acc[c] = v4d{0.0, 0.0, 0.0, 0.0};
  for (int i = 0; i < iters; ++i)
#pragma unroll
    for (int c = 0; c < kCalibChains; ++c) acc[c] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[c], 0, 0, 0);
  double s = 0.0;
#pragma unroll
  for (int c = 0; c < kCalibChains; ++c) s += acc[c].x + acc[c].y + acc[c].z + acc[c].w;
  out[blockIdx.x * 256 + threadIdx.x] = s;
}
}  // namespace

double mfma_issue_rate(int form, int iters, int blocks) {
  double *out = nullptr;
  HN_CHECK(hipMalloc(&out, sizeof(double) * 256 * (size_t)blocks));
  hipEvent_t e0, e1;
  HN_CHECK(hipEventCreate(&e0));
  HN_CHECK(hipEventCreate(&e1));
  auto run = [&] {
    if (form == 0)
      hipLaunchKernelGGL(k_mfma_issue_4x4x4, dim3(blocks), dim3(256), 0, 0, iters, out);
    else
      hipLaunchKernelGGL(k_mfma_issue_16x16x4, dim3(blocks), dim3(256), 0, 0, iters, out);
    HN_CHECK(hipGetLastError());
  };
  run();  // warm-up
  HN_CHECK(hipEventRecord(e0, 0));
  run();
  HN_CHECK(hipEventRecord(e1, 0));
  HN_CHECK(hipEventSynchronize(e1));
  float ms = 0.0f;
  HN_CHECK(hipEventElapsedTime(&ms, e0, e1));
  HN_CHECK(hipEventDestroy(e0));
  HN_CHECK(hipEventDestroy(e1));
  HN_CHECK(hipFree(out));
  const double per_instr = form == 0 ? 512.0 : 2048.0;  // flops per wave instruction
  const double flops = (double)blocks * 4 * iters * kCalibChains * per_instr;
  return flops / (ms * 1e-3) / 1e12;
}

}  // namespace hn
