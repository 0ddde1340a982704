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

int Profiler::begin(int tag, double flops, double bytes, hipStream_t s) {
  std::lock_guard<std::mutex> g(mu_);
  if (!on_) return -1;
  Rec r{tag, flops, bytes, get(), get()};
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

void Profiler::summary(int tag, double *ms, double *flops, double *bytes, long long *launches) {
  std::lock_guard<std::mutex> g(mu_);
  double t = 0.0, f = 0.0, by = 0.0;
  long long n = 0;
  bool f_known = true;
  for (const Rec &r : recs_) {
    if (r.tag != tag) continue;
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

}  // namespace hn
