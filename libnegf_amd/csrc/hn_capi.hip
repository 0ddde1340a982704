// hn_capi.hip -- extern "C" entry points declared in include/hipnegf.h.
//
// Block operations keep the synchronous semantics the reference's Fortran
// relies on (blocking copies, host-returning reductions: gpuroutines.cu:248-996):
// every compute call enqueues on the context's stream and waits for it.
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/hipnegf.h"
#include "hn_engine.h"
#include "hn_gemm.h"
#include "hn_inv.h"
#include "hn_kernels.h"
#include "hn_prof.h"
#include "hn_gen.h"
#include "hn_serve.h"

struct hipnegf_ctx {
  std::unique_ptr<hn::Engine> eng;
};
struct hipnegf_engine_s {
  std::unique_ptr<hn::Engine> eng;
};
struct hipnegf_struct_s {
  std::unique_ptr<hn::DevStruct> st;
};

namespace {

thread_local std::string g_err;

template <typename F> int guard(F &&f) {
  try {
    g_err.clear();
    return f();
  } catch (const std::exception &e) {
    g_err = e.what();
    return -1;
  } catch (...) {
    g_err = "unknown error";
    return -1;
  }
}

constexpr int NT = 256;

template <typename C>
__global__ void k_matsum(int m, int n, C alpha, const C *A, C beta, const C *B, C *Cm,
                         int dagger) {
  long long e = (long long)blockIdx.x * NT + threadIdx.x;
  if (e >= (long long)m * n) return;
  int r = (int)(e % m), c = (int)(e / m);
  C a = (dagger == 1) ? A[c + (long long)r * n] : A[e];
  if (dagger == 1) a.y = -a.y;
  C b = (dagger == 2) ? B[c + (long long)r * n] : B[e];
  if (dagger == 2) b.y = -b.y;
  C out;
  out.x = alpha.x * a.x - alpha.y * a.y + beta.x * b.x - beta.y * b.y;
  out.y = alpha.x * a.y + alpha.y * a.x + beta.x * b.y + beta.y * b.x;
  // in-place aliasing (A or B == C) is safe only when op is N for that operand
  // (gpuroutines.cu:576-577); the element read above happens before the write.
  Cm[e] = out;
}

template <typename C> __global__ void k_initmat(C *A, int n) {
  long long e = (long long)blockIdx.x * NT + threadIdx.x;
  if (e >= (long long)n * n) return;
  int r = (int)(e % n), c = (int)(e / n);
  A[e] = C{(r == c) ? 1.0f : 0.0f, 0.0f};
}

template <typename C, typename R>
__global__ void k_trace(const C *A, int n, const int *mask, R *out) {
  __shared__ double sh[NT];
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += NT)
    if (!mask || mask[i]) s += (double)A[i + (long long)i * n].x;
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int w = NT / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = (R)sh[0];
}

template <typename C, typename R>
__global__ void k_asum(const C *A, long long n, R *out) {
  __shared__ double sh[NT];
  double s = 0.0;
  for (long long i = threadIdx.x; i < n; i += NT) s += fabs((double)A[i].x) + fabs((double)A[i].y);
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int w = NT / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = (R)sh[0];
}

inline unsigned nblk(long long n) { return (unsigned)((n + NT - 1) / NT); }

hipStream_t sof(hipnegf_handle h) { return h && h->eng ? h->eng->stream() : (hipStream_t)0; }

template <typename R> struct CT_;
template <> struct CT_<double> { typedef double2 t; };
template <> struct CT_<float> { typedef float2 t; };

template <typename R>
int multmat(hipnegf_handle h, size_t m, size_t n, size_t k, const void *alpha, const void *A,
            const void *B, const void *beta, void *C, int dagger) {
  typedef typename CT_<R>::t CT;
  if (m > INT32_MAX || n > INT32_MAX || k > INT32_MAX) throw std::runtime_error("size > INT_MAX");
  hn::GemmArgs<R> g;
  g.m = (int)m; g.n = (int)n; g.k = (int)k;
  g.A = (const CT *)A; g.B = (const CT *)B; g.C = (CT *)C; g.ldc = (int)m;
  const CT *al = (const CT *)alpha, *be = (const CT *)beta;
  g.alpha = make_double2(al->x, al->y);
  g.beta = make_double2(be->x, be->y);
  if (dagger == 0) { g.opA = hn::OP_N; g.lda = (int)m; g.opB = hn::OP_N; g.ldb = (int)k; }
  else if (dagger == 1) { g.opA = hn::OP_C; g.lda = (int)k; g.opB = hn::OP_N; g.ldb = (int)k; }
  else if (dagger == 2) { g.opA = hn::OP_N; g.lda = (int)m; g.opB = hn::OP_C; g.ldb = (int)n; }
  else throw std::runtime_error("dagger must be 0, 1 or 2");
  g.batch = 1;
  if constexpr (sizeof(R) == 8) hn::gemm_z(g, sof(h));
  else hn::gemm_c(g, sof(h));
  HN_CHECK(hipStreamSynchronize(sof(h)));
  return 0;
}

template <typename R>
int inverse1(hipnegf_handle h, void *A, void *Ainv, size_t n) {
  typedef typename CT_<R>::t CT;
  if (n == 0) return 0;
  hipStream_t s = sof(h);
  const size_t bytes = n * n * sizeof(CT);
  CT *work = nullptr;
  int *status = nullptr;
  HN_CHECK(hipMallocAsync((void **)&work, bytes, s));
  HN_CHECK(hipMallocAsync((void **)&status, sizeof(int), s));
  HN_CHECK(hipMemcpyAsync(work, A, bytes, hipMemcpyDeviceToDevice, s));
  hn::fill_int(status, 1, 0, s);
  hn::InvWork<R> w;
  void *wb = nullptr;
  if ((int)n > hn::inv_lds_max_n(sizeof(R))) {
    HN_CHECK(hipMallocAsync(&wb, hn::inv_work_bytes((int)n, 1, sizeof(R)), s));
    w = hn::inv_work_carve<R>(wb, (int)n, 1);
  }
  if constexpr (sizeof(R) == 8)
    hn::inv_z((int)n, work, (long long)n * n, (CT *)Ainv, (long long)n * n, 1, nullptr, status, w, s);
  else
    hn::inv_c((int)n, work, (long long)n * n, (CT *)Ainv, (long long)n * n, 1, nullptr, status, w, s);
  int hs = 0;
  HN_CHECK(hipMemcpyAsync(&hs, status, sizeof(int), hipMemcpyDeviceToHost, s));
  HN_CHECK(hipFreeAsync(work, s));
  HN_CHECK(hipFreeAsync(status, s));
  if (wb) HN_CHECK(hipFreeAsync(wb, s));
  HN_CHECK(hipStreamSynchronize(s));
  return hs ? 1 : 0;
}

template <typename R>
int matsum(hipnegf_handle h, size_t m, size_t n, const void *alpha, const void *A,
           const void *beta, const void *B, void *C, int dagger) {
  typedef typename CT_<R>::t CT;
  if (dagger < 0 || dagger > 2) throw std::runtime_error("dagger must be 0, 1 or 2");
  if (A == C && dagger == 1) throw std::runtime_error("matsum: A aliases C with op(A) = A^H");
  if (B == C && dagger == 2) throw std::runtime_error("matsum: B aliases C with op(B) = B^H");
  hipLaunchKernelGGL(k_matsum<CT>, dim3(nblk((long long)m * n)), dim3(NT), 0, sof(h), (int)m,
                     (int)n, *(const CT *)alpha, (const CT *)A, *(const CT *)beta,
                     (const CT *)B, (CT *)C, dagger);
  HN_CHECK(hipGetLastError());
  HN_CHECK(hipStreamSynchronize(sof(h)));
  return 0;
}

template <typename R>
R trace1(hipnegf_handle h, void *A, size_t n, void *h_mask, int mask_present) {
  typedef typename CT_<R>::t CT;
  hipStream_t s = sof(h);
  R *dout = nullptr;
  int *dmask = nullptr;
  HN_CHECK(hipMallocAsync((void **)&dout, sizeof(R), s));
  if (mask_present && h_mask) {
    HN_CHECK(hipMallocAsync((void **)&dmask, n * sizeof(int), s));
    HN_CHECK(hipMemcpyAsync(dmask, h_mask, n * sizeof(int), hipMemcpyHostToDevice, s));
  }
  hipLaunchKernelGGL((k_trace<CT, R>), dim3(1), dim3(NT), 0, s, (const CT *)A, (int)n, dmask, dout);
  HN_CHECK(hipGetLastError());
  R out = 0;
  HN_CHECK(hipMemcpyAsync(&out, dout, sizeof(R), hipMemcpyDeviceToHost, s));
  HN_CHECK(hipFreeAsync(dout, s));
  if (dmask) HN_CHECK(hipFreeAsync(dmask, s));
  HN_CHECK(hipStreamSynchronize(s));
  return out;
}

template <typename R>
int asum1(hipnegf_handle h, void *A, R *sum, size_t n) {
  typedef typename CT_<R>::t CT;
  hipStream_t s = sof(h);
  R *dout = nullptr;
  HN_CHECK(hipMallocAsync((void **)&dout, sizeof(R), s));
  hipLaunchKernelGGL((k_asum<CT, R>), dim3(1), dim3(NT), 0, s, (const CT *)A, (long long)n, dout);
  HN_CHECK(hipGetLastError());
  HN_CHECK(hipMemcpyAsync(sum, dout, sizeof(R), hipMemcpyDeviceToHost, s));
  HN_CHECK(hipFreeAsync(dout, s));
  HN_CHECK(hipStreamSynchronize(s));
  return 0;
}

template <typename R>
int inv_batched(hipnegf_engine e, int n, void *A, long long sA, void *Ainv, long long sAinv,
                       int batch, int *dstatus) {
  typedef typename CT_<R>::t CT;
  HN_CHECK(hipDeviceSynchronize());
  hipStream_t s = e->eng->stream();
  hn::fill_int(dstatus, batch, 0, s);
  hn::InvWork<R> w;
  void *wb = nullptr;
  if (n > hn::inv_lds_max_n(sizeof(R))) {
    HN_CHECK(hipMalloc(&wb, hn::inv_work_bytes(n, batch, sizeof(R))));
    w = hn::inv_work_carve<R>(wb, n, batch);
  }
  if constexpr (sizeof(R) == 8)
    hn::inv_z(n, (CT *)A, sA, (CT *)Ainv, sAinv, batch, nullptr, dstatus, w, s);
  else
    hn::inv_c(n, (CT *)A, sA, (CT *)Ainv, sAinv, batch, nullptr, dstatus, w, s);
  HN_CHECK(hipStreamSynchronize(s));
  if (wb) HN_CHECK(hipFree(wb));
  return 0;
}


// ---- allocation-free block inverse ------------------------------------------
// hipnegf_Zinverse_noalloc keeps its temporaries in `scratch`, one array per
// GPU that lives until the process ends: the working copy of A (the blocked
// inverse overwrites it), the singular flag and the blocked inverse's
// workspace, for one block of order <= kNoAllocMaxN.  Host code never passes
// the array's name to a kernel: the device address of each GPU's copy comes
// from hipGetSymbolAddress, resolved once per device (at hipnegf_blasInit, or
// at the first call on a device) and kept in g_scratch_dev.
constexpr int kNoAllocMaxN = 512;
constexpr size_t kScratchBytes = (size_t)5 << 20;
constexpr int kMaxDevices = 64;
__device__ __align__(256) unsigned char scratch[kScratchBytes];
std::mutex g_scratch_mu;
void *g_scratch_dev[kMaxDevices] = {};

void *scratch_ptr(int dev) {
  if (dev < 0 || dev >= kMaxDevices) throw std::runtime_error("device index out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    int cur = -1;
    HN_CHECK(hipGetDevice(&cur));
    if (cur != dev) throw std::runtime_error("scratch: device " + std::to_string(dev) +
                                             " is not the current device");
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      throw std::runtime_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_dev[dev] = p;
  }
  return g_scratch_dev[dev];
}

// Same kernels and result as inverse1<double>; every temporary in scratch.
int inverse_noalloc(hipnegf_handle h, void *A, void *Ainv, size_t n) {
  if (n == 0) return 0;
  if (!h || !h->eng) throw std::runtime_error("hipnegf_Zinverse_noalloc: null handle");
  if (n > (size_t)kNoAllocMaxN)
    throw std::runtime_error("hipnegf_Zinverse_noalloc: block order " + std::to_string(n) +
                             " exceeds the scratch limit of " + std::to_string(kNoAllocMaxN));
  int cur = -1;
  HN_CHECK(hipGetDevice(&cur));
  if (cur != h->eng->device())
    throw std::runtime_error("hipnegf_Zinverse_noalloc: the current device is not the handle's");
  auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
  const size_t bytes = n * n * sizeof(double2);
  const bool blocked = (int)n > hn::inv_lds_max_n(8);
  const size_t need = al(bytes) + 256 + (blocked ? hn::inv_work_bytes((int)n, 1, 8) : 0);
  if (need > kScratchBytes) throw std::runtime_error("hipnegf_Zinverse_noalloc: scratch too small");
  char *base = (char *)scratch_ptr(cur);
  double2 *work = (double2 *)base;
  int *status = (int *)(base + al(bytes));
  hn::InvWork<double> w;
  if (blocked) w = hn::inv_work_carve<double>(base + al(bytes) + 256, (int)n, 1);
  hipStream_t s = sof(h);
  HN_CHECK(hipMemcpyAsync(work, A, bytes, hipMemcpyDeviceToDevice, s));
  hn::fill_int(status, 1, 0, s);
  hn::inv_z((int)n, work, (long long)n * n, (double2 *)Ainv, (long long)n * n, 1, nullptr, status,
            w, s);
  int hs = 0;
  HN_CHECK(hipMemcpyAsync(&hs, status, sizeof(int), hipMemcpyDeviceToHost, s));
  HN_CHECK(hipStreamSynchronize(s));
  return hs ? 1 : 0;
}

hn::SolveOpts to_opts(const hipnegf_opts *o) {
  hn::SolveOpts so;
  if (o) {
    so.delta = o->delta;
    so.ni = o->ni > 0 ? o->ni : 1;
    so.nf = o->nf > 0 ? o->nf : 2;
    so.precision = o->precision;
    so.max_batch = o->max_batch;
    if (o->ns_steps != 0) so.ns_steps = std::max(0, o->ns_steps);  // -1: none
    if (o->polish > 0) so.polish = o->polish;
    so.eps_drop = o->eps_drop;
    so.share_mirror = (o->flags & HIPNEGF_NO_SHARED_DECIMATION) ? 0 : 1;
    so.graph = (o->flags & HIPNEGF_GRAPH) ? 1 : 0;
    so.fused = (o->flags & HIPNEGF_NO_FUSED) ? 0 : 1;
    so.sgf_mode = (o->flags & HIPNEGF_SGF_LOAD) ? 1 : (o->flags & HIPNEGF_SGF_STORE) ? 2 : 0;
    so.lanes = o->lanes <= 0 ? 0 : std::min(o->lanes, 8);
  }
  return so;
}

}  // namespace

extern "C" {

const char *hipnegf_last_error(void) { return g_err.c_str(); }

// ---- memory ---------------------------------------------------------------
int hipnegf_createMat(void **d, size_t bytes) {
  return guard([&] { HN_CHECK(hipMalloc(d, bytes)); return 0; });
}
int hipnegf_deleteMat(void **d) {
  return guard([&] { HN_CHECK(hipFree(*d)); *d = nullptr; return 0; });
}
int hipnegf_mallocAsync(void **d, size_t bytes) {
  return guard([&] { HN_CHECK(hipMallocAsync(d, bytes, 0)); return 0; });
}
int hipnegf_freeAsync(void **d) {
  return guard([&] { HN_CHECK(hipFreeAsync(*d, 0)); *d = nullptr; return 0; });
}
int hipnegf_copyMatH2D(void *h, void *d, size_t bytes) {
  return guard([&] { HN_CHECK(hipMemcpy(d, h, bytes, hipMemcpyHostToDevice)); return 0; });
}
int hipnegf_copyMatH2D_async(void *h, void *d, size_t bytes) {
  return guard([&] { HN_CHECK(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, 0)); return 0; });
}
int hipnegf_copyMatD2H(void *h, void *d, size_t bytes) {
  return guard([&] { HN_CHECK(hipMemcpy(h, d, bytes, hipMemcpyDeviceToHost)); return 0; });
}
int hipnegf_meminfo(size_t *free_bytes, size_t *total_bytes) {
  return guard([&] { HN_CHECK(hipMemGetInfo(free_bytes, total_bytes)); return 0; });
}
int hipnegf_deviceSynchronize(void) {
  return guard([&] { HN_CHECK(hipDeviceSynchronize()); return 0; });
}

// ---- device / handles -------------------------------------------------------
int hipnegf_getDeviceCount(int *count) {
  return guard([&] { HN_CHECK(hipGetDeviceCount(count)); return 0; });
}
int hipnegf_getDevice(int *device) {
  return guard([&] { HN_CHECK(hipGetDevice(device)); return 0; });
}
int hipnegf_setDevice(int device) {
  return guard([&] { HN_CHECK(hipSetDevice(device)); return 0; });
}
int hipnegf_blasInit(hipnegf_handle *h) {
  return guard([&] {
    int dev = 0;
    HN_CHECK(hipGetDevice(&dev));
    auto *c = new hipnegf_ctx;
    c->eng.reset(new hn::Engine(dev));
    // resolve this device's scratch address once, at set-up; a failure here
    // is reported by the first hipnegf_Zinverse_noalloc call instead
    try { scratch_ptr(dev); } catch (const std::exception &) {}
    *h = c;
    return 0;
  });
}
int hipnegf_blasFinalize(hipnegf_handle h) {
  return guard([&] { delete h; return 0; });
}
int hipnegf_solverInit(hipnegf_handle *h) { return hipnegf_blasInit(h); }
int hipnegf_solverFinalize(hipnegf_handle h) { return hipnegf_blasFinalize(h); }

// ---- compute ----------------------------------------------------------------
int hipnegf_CmultMat(hipnegf_handle h, size_t m, size_t n, size_t k, const void *alpha,
                     const void *A, const void *B, const void *beta, void *C, int dagger) {
  return guard([&] { return multmat<float>(h, m, n, k, alpha, A, B, beta, C, dagger); });
}
int hipnegf_ZmultMat(hipnegf_handle h, size_t m, size_t n, size_t k, const void *alpha,
                     const void *A, const void *B, const void *beta, void *C, int dagger) {
  return guard([&] { return multmat<double>(h, m, n, k, alpha, A, B, beta, C, dagger); });
}
int hipnegf_Cinverse(hipnegf_handle hb, hipnegf_handle, void *A, void *Ainv, size_t n) {
  return guard([&] { return inverse1<float>(hb, A, Ainv, n); });
}
int hipnegf_Zinverse(hipnegf_handle hb, hipnegf_handle, void *A, void *Ainv, size_t n) {
  return guard([&] { return inverse1<double>(hb, A, Ainv, n); });
}
int hipnegf_Zinverse_noalloc(hipnegf_handle hb, void *A, void *Ainv, size_t n) {
  return guard([&] { return inverse_noalloc(hb, A, Ainv, n); });
}
int hipnegf_Cmatsum(hipnegf_handle h, size_t m, size_t n, const void *alpha, const void *A,
                    const void *beta, const void *B, void *C, int dagger) {
  return guard([&] { return matsum<float>(h, m, n, alpha, A, beta, B, C, dagger); });
}
int hipnegf_Zmatsum(hipnegf_handle h, size_t m, size_t n, const void *alpha, const void *A,
                    const void *beta, const void *B, void *C, int dagger) {
  return guard([&] { return matsum<double>(h, m, n, alpha, A, beta, B, C, dagger); });
}
int hipnegf_Cinitmat(void *A, size_t n) {
  return guard([&] {
    hipLaunchKernelGGL(k_initmat<float2>, dim3(nblk((long long)n * n)), dim3(NT), 0, 0, (float2 *)A, (int)n);
    HN_CHECK(hipGetLastError());
    HN_CHECK(hipStreamSynchronize(0));
    return 0;
  });
}
int hipnegf_Zinitmat(void *A, size_t n) {
  return guard([&] {
    hipLaunchKernelGGL(k_initmat<double2>, dim3(nblk((long long)n * n)), dim3(NT), 0, 0, (double2 *)A, (int)n);
    HN_CHECK(hipGetLastError());
    HN_CHECK(hipStreamSynchronize(0));
    return 0;
  });
}
float hipnegf_Ctrace(hipnegf_handle h, void *A, size_t n, void *h_mask, int mask_present) {
  float r = 0.0f;
  guard([&] { r = trace1<float>(h, A, n, h_mask, mask_present); return 0; });
  return r;
}
double hipnegf_Ztrace(hipnegf_handle h, void *A, size_t n, void *h_mask, int mask_present) {
  double r = 0.0;
  guard([&] { r = trace1<double>(h, A, n, h_mask, mask_present); return 0; });
  return r;
}
int hipnegf_Cmatcopy(hipnegf_handle h, const void *A, void *B, size_t nelem) {
  return guard([&] {
    HN_CHECK(hipMemcpyAsync(B, A, nelem * sizeof(float2), hipMemcpyDeviceToDevice, sof(h)));
    HN_CHECK(hipStreamSynchronize(sof(h)));
    return 0;
  });
}
int hipnegf_Zmatcopy(hipnegf_handle h, const void *A, void *B, size_t nelem) {
  return guard([&] {
    HN_CHECK(hipMemcpyAsync(B, A, nelem * sizeof(double2), hipMemcpyDeviceToDevice, sof(h)));
    HN_CHECK(hipStreamSynchronize(sof(h)));
    return 0;
  });
}
int hipnegf_Casum(hipnegf_handle h, void *A, float *sum, size_t n) {
  return guard([&] { return asum1<float>(h, A, sum, n); });
}
int hipnegf_Zasum(hipnegf_handle h, void *A, double *sum, size_t n) {
  return guard([&] { return asum1<double>(h, A, sum, n); });
}

int hipnegf_Zdecimation(hipnegf_handle hb, hipnegf_handle, void *h_Go_out, void *h_Ao,
                        void *h_Bo, void *h_Co, size_t n, int tf32, int *ncyc, double SGFACC) {
  (void)tf32;
  return guard([&] {
    int st = hb->eng->decimation_host((int)n, (double2 *)h_Go_out, (const double2 *)h_Ao,
                                      (const double2 *)h_Bo, (const double2 *)h_Co, SGFACC, 1, ncyc);
    return (st & hn::ST_NOT_CONVERGED) ? 2 : (st ? 1 : 0);
  });
}
int hipnegf_Cdecimation(hipnegf_handle hb, hipnegf_handle, void *h_Go_out, void *h_Ao,
                        void *h_Bo, void *h_Co, size_t n, int tf32, int *ncyc, float SGFACC) {
  (void)tf32;
  return guard([&] {
    // complex-float host data: widen, run the FP64 iteration, narrow the result
    size_t n2 = n * n;
    std::vector<double2> a(n2), b(n2), c(n2), g(n2);
    const float2 *fa = (const float2 *)h_Ao, *fb = (const float2 *)h_Bo, *fc = (const float2 *)h_Co;
    for (size_t i = 0; i < n2; ++i) {
      a[i] = make_double2(fa[i].x, fa[i].y);
      b[i] = make_double2(fb[i].x, fb[i].y);
      c[i] = make_double2(fc[i].x, fc[i].y);
    }
    int st = hb->eng->decimation_host((int)n, g.data(), a.data(), b.data(), c.data(), SGFACC, 1, ncyc);
    float2 *go = (float2 *)h_Go_out;
    for (size_t i = 0; i < n2; ++i) go[i] = make_float2((float)g[i].x, (float)g[i].y);
    return (st & hn::ST_NOT_CONVERGED) ? 2 : (st ? 1 : 0);
  });
}

// ---- batched engine ------------------------------------------------------------
int hipnegf_engine_create(int device, hipnegf_engine *e) {
  return guard([&] {
    auto *x = new hipnegf_engine_s;
    x->eng.reset(new hn::Engine(device));
    *e = x;
    return 0;
  });
}
int hipnegf_engine_destroy(hipnegf_engine e) {
  return guard([&] { delete e; return 0; });
}
int hipnegf_engine_sync(hipnegf_engine e) {
  return guard([&] { HN_CHECK(hipStreamSynchronize(e->eng->stream())); return 0; });
}
void *hipnegf_engine_stream(hipnegf_engine e) { return (void *)e->eng->stream(); }

int hipnegf_struct_create(int nbl, const int *bsz, int ncont, const int *cblk, const int *nc,
                          int smode, int nslots, hipnegf_struct *st) {
  return guard([&] {
    std::unique_ptr<hipnegf_struct_s> x(new hipnegf_struct_s);
    x->st.reset(new hn::DevStruct(nbl, bsz, ncont, cblk, nc, smode, nslots));
    *st = x.release();
    return 0;
  });
}
int hipnegf_struct_create_surf(int nbl, const int *bsz, int ncont, const int *cblk,
                               const int *nc, const int *ns, int smode, int nslots,
                               hipnegf_struct *st) {
  return guard([&] {
    std::unique_ptr<hipnegf_struct_s> x(new hipnegf_struct_s);
    x->st.reset(new hn::DevStruct(nbl, bsz, ncont, cblk, nc, smode, nslots, ns));
    *st = x.release();
    return 0;
  });
}
int hipnegf_struct_destroy(hipnegf_struct st) {
  return guard([&] { delete st; return 0; });
}

double2 *block_ptr(hn::DevStruct &s, int slot, int kind, int index, long long *n) {
  if (slot < 0 || slot >= s.nslots) throw std::runtime_error("slot out of range");
  if (kind < 0 || kind >= hn::BK_COUNT) throw std::runtime_error("bad block kind");
  if (s.off[kind].empty()) throw std::runtime_error("block kind not present (smode)");
  if (index < 0 || index >= (int)s.off[kind].size()) throw std::runtime_error("block index out of range");
  *n = s.block_elems(kind, index);
  return s.pool + (long long)slot * s.slot_elems + s.off[kind][index];
}

int hipnegf_struct_set_block(hipnegf_struct st, int slot, int kind, int index, const double *host) {
  return guard([&] {
    long long n = 0;
    double2 *d = block_ptr(*st->st, slot, kind, index, &n);
    HN_CHECK(hipMemcpy(d, host, n * sizeof(double2), hipMemcpyHostToDevice));
    st->st->touched();
    return 0;
  });
}
int hipnegf_struct_set_block_device(hipnegf_struct st, int slot, int kind, int index,
                                    const void *dev) {
  return guard([&] {
    long long n = 0;
    double2 *d = block_ptr(*st->st, slot, kind, index, &n);
    HN_CHECK(hipMemcpy(d, dev, n * sizeof(double2), hipMemcpyDeviceToDevice));
    st->st->touched();
    return 0;
  });
}

int hipnegf_solve_transmission_device(hipnegf_engine e, hipnegf_struct st, int nitems,
                                      const double *dE, const int *dslot,
                                      const hipnegf_opts *o, double *dT, int *dncyc,
                                      int *dstatus) {
  return guard([&] {
    e->eng->transmission(*st->st, nitems, dE, dslot, to_opts(o), dT, dncyc, dstatus);
    return 0;
  });
}

int hipnegf_solve_transmission(hipnegf_engine e, hipnegf_struct st, int nitems,
                               const double *E, const int *slot, const hipnegf_opts *o,
                               double *T, int *ncyc, int *status) {
  return guard([&] {
    if (nitems <= 0) return 0;
    for (int i = 0; i < nitems; ++i)
      if (slot[i] < 0 || slot[i] >= st->st->nslots) throw std::runtime_error("item slot out of range");
    const int nc = std::max(st->st->ncont, 1);
    double *dE = nullptr, *dT = nullptr;
    int *ds = nullptr, *dn = nullptr, *dst = nullptr;
    HN_CHECK(hipMalloc(&dE, sizeof(double) * nitems));
    HN_CHECK(hipMalloc(&dT, sizeof(double) * nitems));
    HN_CHECK(hipMalloc(&ds, sizeof(int) * nitems));
    HN_CHECK(hipMalloc(&dn, sizeof(int) * nitems * nc));
    HN_CHECK(hipMalloc(&dst, sizeof(int) * nitems));
    struct F { void *p[5]; ~F() { for (void *q : p) if (q) (void)hipFree(q); } } fr{{dE, dT, ds, dn, dst}};
    HN_CHECK(hipMemcpy(dE, E, sizeof(double) * nitems, hipMemcpyHostToDevice));
    HN_CHECK(hipMemcpy(ds, slot, sizeof(int) * nitems, hipMemcpyHostToDevice));
    e->eng->transmission(*st->st, nitems, dE, ds, to_opts(o), dT, dn, dst);
    HN_CHECK(hipMemcpy(T, dT, sizeof(double) * nitems, hipMemcpyDeviceToHost));
    if (ncyc) HN_CHECK(hipMemcpy(ncyc, dn, sizeof(int) * nitems * nc, hipMemcpyDeviceToHost));
    if (status) HN_CHECK(hipMemcpy(status, dst, sizeof(int) * nitems, hipMemcpyDeviceToHost));
    return 0;
  });
}

int hipnegf_solve_transmission_pairs(hipnegf_engine e, hipnegf_struct st, int nitems,
                                     const double *E, const int *slot, const hipnegf_opts *o,
                                     int npairs, const int *ni, const int *nf, double *T,
                                     int *ncyc, int *status) {
  return guard([&] {
    if (nitems <= 0) return 0;
    if (npairs < 1) throw std::runtime_error("npairs < 1");
    for (int i = 0; i < nitems; ++i)
      if (slot[i] < 0 || slot[i] >= st->st->nslots) throw std::runtime_error("item slot out of range");
    const int nc = std::max(st->st->ncont, 1);
    void *p[5] = {};
    struct F { void **p; ~F() { for (int i = 0; i < 5; ++i) if (p[i]) (void)hipFree(p[i]); } } fr{p};
    HN_CHECK(hipMalloc(&p[0], sizeof(double) * nitems));
    HN_CHECK(hipMalloc(&p[1], sizeof(double) * nitems * npairs));
    HN_CHECK(hipMalloc(&p[2], sizeof(int) * nitems));
    HN_CHECK(hipMalloc(&p[3], sizeof(int) * nitems * nc));
    HN_CHECK(hipMalloc(&p[4], sizeof(int) * nitems));
    HN_CHECK(hipMemcpy(p[0], E, sizeof(double) * nitems, hipMemcpyHostToDevice));
    HN_CHECK(hipMemcpy(p[2], slot, sizeof(int) * nitems, hipMemcpyHostToDevice));
    e->eng->transmission_pairs(*st->st, nitems, (const double *)p[0], (const int *)p[2], to_opts(o),
                               npairs, ni, nf, (double *)p[1], (int *)p[3], (int *)p[4]);
    HN_CHECK(hipMemcpy(T, p[1], sizeof(double) * nitems * npairs, hipMemcpyDeviceToHost));
    if (ncyc) HN_CHECK(hipMemcpy(ncyc, p[3], sizeof(int) * nitems * nc, hipMemcpyDeviceToHost));
    if (status) HN_CHECK(hipMemcpy(status, p[4], sizeof(int) * nitems, hipMemcpyDeviceToHost));
    return 0;
  });
}

int hipnegf_solve_transmission_dos(hipnegf_engine e, hipnegf_struct st, int nitems,
                                   const double *E, const int *slot, const hipnegf_opts *o,
                                   int npairs, const int *ni, const int *nf,
                                   const hipnegf_dos_opts *d, double *T, double *ldos, int *ncyc,
                                   int *status) {
  return guard([&] {
    if (nitems <= 0) return 0;
    if (npairs < 1) throw std::runtime_error("npairs < 1");
    if (!d) throw std::runtime_error("LDOS options missing");
    hn::DevStruct &S = *st->st;
    for (int i = 0; i < nitems; ++i)
      if (slot[i] < 0 || slot[i] >= S.nslots) throw std::runtime_error("item slot out of range");
    const int nl = d->full ? 0 : d->nldos;
    if (nl < 0) throw std::runtime_error("nldos < 0");
    if (nl > 0) {
      if (!d->ptr || !d->idx) throw std::runtime_error("LDOS projection lists missing");
      if (d->ptr[0] != 0) throw std::runtime_error("LDOS ptr[0] must be 0");
      for (int l = 0; l < nl; ++l)
        if (d->ptr[l + 1] < d->ptr[l]) throw std::runtime_error("LDOS ptr not ascending");
      for (int k = 0; k < d->ptr[nl]; ++k)
        if (d->idx[k] < 1) throw std::runtime_error("LDOS orbital index < 1");
    }
    const long long nout = d->full ? S.ndev() : nl;
    const int nc = std::max(S.ncont, 1);
    const int nidx = nl > 0 ? d->ptr[nl] : 0;
    void *p[7] = {};
    struct F { void **p; ~F() { for (int i = 0; i < 7; ++i) if (p[i]) (void)hipFree(p[i]); } } fr{p};
    HN_CHECK(hipMalloc(&p[0], sizeof(double) * nitems));
    HN_CHECK(hipMalloc(&p[1], sizeof(double) * nitems * npairs));
    HN_CHECK(hipMalloc(&p[2], sizeof(int) * nitems));
    HN_CHECK(hipMalloc(&p[3], sizeof(int) * nitems * nc));
    HN_CHECK(hipMalloc(&p[4], sizeof(int) * nitems));
    HN_CHECK(hipMalloc(&p[5], sizeof(double) * std::max<long long>(nitems * nout, 1)));
    HN_CHECK(hipMalloc(&p[6], sizeof(int) * (nl + 1 + std::max(nidx, 1))));
    HN_CHECK(hipMemcpy(p[0], E, sizeof(double) * nitems, hipMemcpyHostToDevice));
    HN_CHECK(hipMemcpy(p[2], slot, sizeof(int) * nitems, hipMemcpyHostToDevice));
    hn::DosOpts dd;
    dd.dos_delta = d->dos_delta;
    dd.full = d->full ? 1 : 0;
    dd.nldos = nl;
    if (nl > 0) {
      HN_CHECK(hipMemcpy(p[6], d->ptr, sizeof(int) * (nl + 1), hipMemcpyHostToDevice));
      HN_CHECK(hipMemcpy((int *)p[6] + nl + 1, d->idx, sizeof(int) * nidx, hipMemcpyHostToDevice));
      dd.ptr = (const int *)p[6];
      dd.idx = (const int *)p[6] + nl + 1;
    }
    e->eng->transmission_dos(S, nitems, (const double *)p[0], (const int *)p[2], to_opts(o), npairs,
                             ni, nf, dd, (double *)p[1], nout > 0 ? (double *)p[5] : nullptr,
                             (int *)p[3], (int *)p[4]);
    HN_CHECK(hipMemcpy(T, p[1], sizeof(double) * nitems * npairs, hipMemcpyDeviceToHost));
    if (ldos && nout > 0)
      HN_CHECK(hipMemcpy(ldos, p[5], sizeof(double) * nitems * nout, hipMemcpyDeviceToHost));
    if (ncyc) HN_CHECK(hipMemcpy(ncyc, p[3], sizeof(int) * nitems * nc, hipMemcpyDeviceToHost));
    if (status) HN_CHECK(hipMemcpy(status, p[4], sizeof(int) * nitems, hipMemcpyDeviceToHost));
    return 0;
  });
}

int hipnegf_solve_scba_mw(hipnegf_engine e, hipnegf_struct st, int nitems, const double *E,
                          const int *slot, const hipnegf_opts *o, const double *coupling,
                          int niter, double tol, const double *frm, int frm_per_item, int ref,
                          double *curr, int *iters, double *err, int *status, double *gn_diag,
                          double *gr_diag) {
  return guard([&] {
    if (nitems <= 0) return 0;
    hn::DevStruct &S = *st->st;
    for (int i = 0; i < nitems; ++i)
      if (slot[i] < 0 || slot[i] >= S.nslots) throw std::runtime_error("item slot out of range");
    if (ref < 1 || ref > S.ncont + 1) throw std::runtime_error("reference contact out of range");
    if (niter < 0) throw std::runtime_error("niter < 0");
    int ndev = 0;
    for (int b : S.bsz) ndev += b;
    const int nc = std::max(S.ncont, 1);
    const long long nf = (long long)(frm_per_item ? nitems : 1) * (S.ncont + 1);
    std::vector<double> zero_c;
    if (!coupling) zero_c.assign(ndev, 0.0);
    void *p[11] = {};
    struct F { void **p; ~F() { for (int i = 0; i < 11; ++i) if (p[i]) (void)hipFree(p[i]); } } fr{p};
    HN_CHECK(hipMalloc(&p[0], sizeof(double) * nitems));            // E
    HN_CHECK(hipMalloc(&p[1], sizeof(int) * nitems));               // slot
    HN_CHECK(hipMalloc(&p[2], sizeof(double) * ndev));              // coupling
    HN_CHECK(hipMalloc(&p[3], sizeof(double) * nf));                // frm
    HN_CHECK(hipMalloc(&p[4], sizeof(double) * nitems * nc));       // curr
    HN_CHECK(hipMalloc(&p[5], sizeof(int) * nitems));               // iters
    HN_CHECK(hipMalloc(&p[6], sizeof(double) * nitems));            // err
    HN_CHECK(hipMalloc(&p[7], sizeof(int) * nitems));               // status
    if (gn_diag) HN_CHECK(hipMalloc(&p[8], sizeof(double2) * (size_t)nitems * ndev));
    if (gr_diag) HN_CHECK(hipMalloc(&p[9], sizeof(double2) * (size_t)nitems * ndev));
    HN_CHECK(hipMemcpy(p[0], E, sizeof(double) * nitems, hipMemcpyHostToDevice));
    HN_CHECK(hipMemcpy(p[1], slot, sizeof(int) * nitems, hipMemcpyHostToDevice));
    HN_CHECK(hipMemcpy(p[2], coupling ? coupling : zero_c.data(), sizeof(double) * ndev,
                       hipMemcpyHostToDevice));
    HN_CHECK(hipMemcpy(p[3], frm, sizeof(double) * nf, hipMemcpyHostToDevice));
    hn::SCBAOpts so;
    so.coupling = (const double *)p[2];
    so.niter = niter;
    so.tol = tol;
    so.frm = (const double *)p[3];
    so.frm_per_item = frm_per_item ? 1 : 0;
    so.ref = ref;
    e->eng->scba_mw(S, nitems, (const double *)p[0], (const int *)p[1], to_opts(o), so, (double *)p[4],
                    (int *)p[5], (double *)p[6], (int *)p[7], (double2 *)p[8], (double2 *)p[9]);
    HN_CHECK(hipMemcpy(curr, p[4], sizeof(double) * nitems * S.ncont, hipMemcpyDeviceToHost));
    if (iters) HN_CHECK(hipMemcpy(iters, p[5], sizeof(int) * nitems, hipMemcpyDeviceToHost));
    if (err) HN_CHECK(hipMemcpy(err, p[6], sizeof(double) * nitems, hipMemcpyDeviceToHost));
    if (status) HN_CHECK(hipMemcpy(status, p[7], sizeof(int) * nitems, hipMemcpyDeviceToHost));
    if (gn_diag) HN_CHECK(hipMemcpy(gn_diag, p[8], sizeof(double2) * (size_t)nitems * ndev, hipMemcpyDeviceToHost));
    if (gr_diag) HN_CHECK(hipMemcpy(gr_diag, p[9], sizeof(double2) * (size_t)nitems * ndev, hipMemcpyDeviceToHost));
    return 0;
  });
}

int hipnegf_solve_inelastic_mw(hipnegf_engine e, hipnegf_struct st, int ne, const double *E,
                               int nk, const double *kweights, const hipnegf_opts *o,
                               const double *coupling, int niter_ela, double tol_ela,
                               const double *frm, int ref, const double *K, int ndz, const int *iz,
                               int shift, double Nq, int tridiag, int niter, double tol,
                               int integration, double g_spin, double *curr_mat, double *currents,
                               int *passes, double *err, int *status, int *iters) {
  return guard([&] {
    hn::DevStruct &S = *st->st;
    if (ne < 2) throw std::runtime_error("inelastic SCBA needs a grid of at least 2 energies");
    if (nk < 1 || nk > S.nslots) throw std::runtime_error("k-point count out of range");
    if (ref < 1 || ref > S.ncont + 1) throw std::runtime_error("reference contact out of range");
    if (niter < 0 || niter_ela < 0) throw std::runtime_error("niter < 0");
    const double dE = E[1] - E[0];
    for (int i = 1; i < ne; ++i)
      if (std::fabs((E[i] - E[i - 1]) - dE) > 1e-9 * std::max(1.0, std::fabs(dE)))
        throw std::runtime_error("inelastic SCBA needs a uniform energy grid");
    int ndev = 0;
    for (int b : S.bsz) ndev += b;
    void *pc = nullptr;
    struct F { void **p; ~F() { if (*p) (void)hipFree(*p); } } fr{&pc};
    if (coupling) {
      HN_CHECK(hipMalloc(&pc, sizeof(double) * ndev));
      HN_CHECK(hipMemcpy(pc, coupling, sizeof(double) * ndev, hipMemcpyHostToDevice));
    }
    hn::SCBAOpts so;
    so.coupling = (const double *)pc;
    so.niter = niter_ela;
    so.tol = tol_ela;
    so.ref = ref;
    hn::InelOpts io;
    io.ne = ne; io.nk = nk; io.kw = kweights; io.K = K; io.ndz = ndz; io.iz = iz;
    io.shift = shift; io.Nq = Nq; io.tridiag = tridiag ? 1 : 0; io.niter = niter; io.tol = tol;
    io.integration = integration; io.estep = dE; io.g_spin = g_spin;
    int np = 0;
    double er = 0.0;
    e->eng->inel_mw(S, E, frm, to_opts(o), so, io, curr_mat, currents, &np, &er, status, iters);
    if (passes) *passes = np;
    if (err) *err = er;
    return 0;
  });
}

// set_reciprocal_vectors (elphinel.F90:1118-1143) and the absolute k-points
// of prepare_POKmat / prepare_NonPOKmat (:345-350): rec . kfrac, or 0 when
// every lattice vector is 0
static void inel_kabs(int nk, const double *kfrac, const double *latt, std::vector<double> &ka) {
  ka.assign(3 * (size_t)nk, 0.0);
  bool zero = true;
  for (int t = 0; t < 9; ++t) zero = zero && latt[t] == 0.0;
  if (zero) return;
  const double *a1 = latt, *a2 = latt + 3, *a3 = latt + 6;
  auto cross = [](const double *u, const double *v, double *w) {
    w[0] = u[1] * v[2] - u[2] * v[1];
    w[1] = u[2] * v[0] - u[0] * v[2];
    w[2] = u[0] * v[1] - u[1] * v[0];
  };
  double c[3][3];
  cross(a2, a3, c[0]);
  cross(a3, a1, c[1]);
  cross(a1, a2, c[2]);
  const double vol = a1[0] * c[0][0] + a1[1] * c[0][1] + a1[2] * c[0][2];
  const double tp = 2.0 * 3.14159265358979323846;
  for (int k = 0; k < nk; ++k)
    for (int r = 0; r < 3; ++r) {
      double v = 0.0;
      for (int j = 0; j < 3; ++j) v += tp * c[j][r] / vol * kfrac[j + 3 * k];
      ka[r + 3 * (size_t)k] = v;
    }
}

int hipnegf_inel_kmat_po(int ndz, double dz, int nk, const double *kfrac, const double *kweights,
                         const double *latt, double coupling, double wq, double cell_area,
                         double eps0, double eps_inf, double q0, double *K) {
  return guard([&] {
    if (ndz < 1 || nk < 1) throw std::runtime_error("empty K table");
    std::vector<double> ka;
    inel_kabs(nk, kfrac, latt, ka);
    const double Ce = coupling * wq / 2.0 / cell_area * (1.0 / eps_inf - 1.0 / eps0);
    for (int iQ = 0; iQ < nk; ++iQ)
      for (int iK = 0; iK < nk; ++iK) {
        double Q2 = 0.0;
        for (int r = 0; r < 3; ++r) {
          const double d = ka[r + 3 * (size_t)iK] - ka[r + 3 * (size_t)iQ];
          Q2 += d * d;
        }
        const double bb = std::sqrt(q0 * q0 + Q2);
        for (int z = 0; z < ndz; ++z) {
          const double zz = z * dz;
          const double Kf = (2.0 * Q2 + q0 * q0 * (1.0 - bb * zz)) * std::exp(-bb * zz) / (4.0 * bb * bb * bb);
          K[z + (size_t)ndz * (iK + (size_t)nk * iQ)] = Ce * kweights[iQ] * Kf;
        }
      }
    return 0;
  });
}

int hipnegf_inel_kmat_nonpo(int ndz, double dz, int nk, const double *kfrac,
                            const double *kweights, const double *latt, double coupling, double D0,
                            double *K) {
  return guard([&] {
    if (ndz < 1 || nk < 1) throw std::runtime_error("empty K table");
    std::vector<double> ka;
    inel_kabs(nk, kfrac, latt, ka);
    for (int iQ = 0; iQ < nk; ++iQ) {
      const double *kq = &ka[3 * (size_t)iQ];
      const double kq2 = kq[0] * kq[0] + kq[1] * kq[1] + kq[2] * kq[2];
      for (int iK = 0; iK < nk; ++iK) {
        const double *kk = &ka[3 * (size_t)iK];
        const double kk2 = kk[0] * kk[0] + kk[1] * kk[1] + kk[2] * kk[2];
        const double QQ = kk[0] * kq[0] + kk[1] * kq[1] + kk[2] * kq[2];
        for (int z = 0; z < ndz; ++z) {
          double v = 0.0;
          if (kq2 != 0.0 && kk2 != 0.0)
            v = coupling * kweights[iQ] * (D0 * QQ * QQ * std::exp(-std::sqrt(kq2) * (z * dz)) / (2.0 * kq2 * kk2));
          K[z + (size_t)ndz * (iK + (size_t)nk * iQ)] = v;
        }
      }
    }
    return 0;
  });
}

int hipnegf_zgemm_batched(hipnegf_engine e, int m, int n, int k, const double *alpha,
                          const void *A, long long sA, int lda, int opA, const void *B,
                          long long sB, int ldb, int opB, const double *beta, void *C,
                          long long sC, int ldc, int batch) {
  return guard([&] {
    HN_CHECK(hipDeviceSynchronize());
    hn::GemmArgs<double> g;
    g.m = m; g.n = n; g.k = k;
    g.A = (const double2 *)A; g.sA = sA; g.lda = lda; g.opA = opA;
    g.B = (const double2 *)B; g.sB = sB; g.ldb = ldb; g.opB = opB;
    g.C = (double2 *)C; g.sC = sC; g.ldc = ldc;
    g.alpha = make_double2(alpha[0], alpha[1]);
    g.beta = make_double2(beta[0], beta[1]);
    g.batch = batch;
    hn::gemm_z(g, e->eng->stream());
    HN_CHECK(hipStreamSynchronize(e->eng->stream()));
    return 0;
  });
}
int hipnegf_cgemm_batched(hipnegf_engine e, int m, int n, int k, const double *alpha,
                          const void *A, long long sA, int lda, int opA, const void *B,
                          long long sB, int ldb, int opB, const double *beta, void *C,
                          long long sC, int ldc, int batch) {
  return guard([&] {
    HN_CHECK(hipDeviceSynchronize());
    hn::GemmArgs<float> g;
    g.m = m; g.n = n; g.k = k;
    g.A = (const float2 *)A; g.sA = sA; g.lda = lda; g.opA = opA;
    g.B = (const float2 *)B; g.sB = sB; g.ldb = ldb; g.opB = opB;
    g.C = (float2 *)C; g.sC = sC; g.ldc = ldc;
    g.alpha = make_double2(alpha[0], alpha[1]);
    g.beta = make_double2(beta[0], beta[1]);
    g.batch = batch;
    hn::gemm_c(g, e->eng->stream());
    HN_CHECK(hipStreamSynchronize(e->eng->stream()));
    return 0;
  });
}

int hipnegf_zinverse_batched(hipnegf_engine e, int n, void *A, long long sA, void *Ainv,
                             long long sAinv, int batch, int *dstatus) {
  return guard([&] { return inv_batched<double>(e, n, A, sA, Ainv, sAinv, batch, dstatus); });
}
int hipnegf_cinverse_batched(hipnegf_engine e, int n, void *A, long long sA, void *Ainv,
                             long long sAinv, int batch, int *dstatus) {
  return guard([&] { return inv_batched<float>(e, n, A, sA, Ainv, sAinv, batch, dstatus); });
}

int hipnegf_engine_set_sgf_cache(hipnegf_engine e, long long bytes) {
  return guard([&] {
    if (bytes < 0) throw std::runtime_error("negative cache size");
    e->eng->set_sgf_cache((size_t)bytes);
    return 0;
  });
}
int hipnegf_engine_sgf_stats(hipnegf_engine e, long long *hits, long long *stores) {
  return guard([&] {
    if (hits) *hits = e->eng->sgf_hits();
    if (stores) *stores = e->eng->sgf_stores();
    return 0;
  });
}
int hipnegf_struct_set_fictitious(hipnegf_struct st, int contact, int fictitious, double dos) {
  return guard([&] {
    hn::DevStruct &S = *st->st;
    if (contact < 0 || contact >= S.ncont) throw std::runtime_error("contact out of range");
    S.fict[contact] = fictitious ? 1 : 0;
    S.contact_dos[contact] = dos;
    S.touched();
    return 0;
  });
}

int hipnegf_density_layout(hipnegf_struct st, long long *offsets) {
  return guard([&] {
    const std::vector<long long> lay = hn::Engine::density_layout(*st->st);
    for (size_t i = 0; i < lay.size(); ++i) offsets[i] = lay[i];
    return 0;
  });
}

int hipnegf_density_accumulate(hipnegf_engine e, hipnegf_struct st, int nitems, const double *Er,
                               const double *Ei, const int *slot, const hipnegf_opts *o, int mode,
                               const double *w, const double *frm, int ref, int outer,
                               double *acc, int *status) {
  return guard([&] {
    if (nitems <= 0) return 0;
    hn::DevStruct &S = *st->st;
    for (int i = 0; i < nitems; ++i)
      if (slot[i] < 0 || slot[i] >= S.nslots) throw std::runtime_error("item slot out of range");
    if (mode != 0 && mode != 1) throw std::runtime_error("mode must be 0 (contour) or 1 (real axis)");
    if (outer < 0 || outer > 2) throw std::runtime_error("outer must be 0, 1 or 2");
    if (mode == 0 && !Ei) throw std::runtime_error("contour points need imaginary parts");
    if (mode == 1 && (!frm || ref < 1 || ref > S.ncont + 1))
      throw std::runtime_error("real-axis points need occupations and a reference contact");
    const long long tot = hn::Engine::density_layout(S).back();
    const int nf = S.ncont + 1;
    void *p[6] = {};
    struct F { void **p; ~F() { for (int i = 0; i < 6; ++i) if (p[i]) (void)hipFree(p[i]); } } fr{p};
    HN_CHECK(hipMalloc(&p[0], sizeof(double) * nitems * 2));
    HN_CHECK(hipMalloc(&p[1], sizeof(int) * nitems * 2));
    HN_CHECK(hipMalloc(&p[2], sizeof(double) * nitems * 2));
    HN_CHECK(hipMalloc(&p[3], sizeof(double) * std::max(nitems * nf, 1)));
    HN_CHECK(hipMalloc(&p[4], sizeof(double) * 2 * std::max<long long>(tot, 1)));
    double *dE = (double *)p[0], *dEi = dE + nitems;
    int *dsl = (int *)p[1], *dst = dsl + nitems;
    HN_CHECK(hipMemcpy(dE, Er, sizeof(double) * nitems, hipMemcpyHostToDevice));
    if (mode == 0) HN_CHECK(hipMemcpy(dEi, Ei, sizeof(double) * nitems, hipMemcpyHostToDevice));
    HN_CHECK(hipMemcpy(dsl, slot, sizeof(int) * nitems, hipMemcpyHostToDevice));
    HN_CHECK(hipMemcpy(p[2], w, sizeof(double) * 2 * nitems, hipMemcpyHostToDevice));
    if (mode == 1) HN_CHECK(hipMemcpy(p[3], frm, sizeof(double) * nitems * nf, hipMemcpyHostToDevice));
    HN_CHECK(hipMemcpy(p[4], acc, sizeof(double) * 2 * tot, hipMemcpyHostToDevice));
    hn::DensityOpts d;
    d.mode = mode;
    d.w = (const double2 *)p[2];
    d.frm = mode == 1 ? (const double *)p[3] : nullptr;
    d.ref = ref;
    d.outer = outer;
    d.status = dst;
    e->eng->density(S, nitems, dE, mode == 0 ? dEi : nullptr, dsl, to_opts(o), d, (double2 *)p[4]);
    HN_CHECK(hipMemcpy(acc, p[4], sizeof(double) * 2 * tot, hipMemcpyDeviceToHost));
    if (status) HN_CHECK(hipMemcpy(status, dst, sizeof(int) * nitems, hipMemcpyDeviceToHost));
    return 0;
  });
}

int hipnegf_struct_set_tun_proj(hipnegf_struct st, int n, const int *idx) {
  return guard([&] {
    if (n > 0 && !idx) throw std::runtime_error("tun_proj indexes missing");
    st->st->set_tun_proj(n, idx);
    return 0;
  });
}

int hipnegf_profile_enable(int on) {
  return guard([&] { hn::profiler().enable(on != 0); return 0; });
}
int hipnegf_profile_read(int tag, double *ms, double *flops, double *bytes, long long *launches) {
  return guard([&] {
    if (tag < 0 || tag >= hn::PT_COUNT) throw std::runtime_error("bad profile tag");
    hn::profiler().summary(tag, ms, flops, bytes, launches);
    return 0;
  });
}
int hipnegf_profile_read_variant(int tag, int variant, double *ms, double *flops, double *bytes,
                                 long long *launches) {
  return guard([&] {
    if (tag < 0 || tag >= hn::PT_COUNT) throw std::runtime_error("bad profile tag");
    if (variant < 0 || variant > 3) throw std::runtime_error("bad kernel variant");
    hn::profiler().summary(tag, ms, flops, bytes, launches, variant);
    return 0;
  });
}
int hipnegf_profile_clear(void) {
  return guard([&] { hn::profiler().clear(); return 0; });
}
int hipnegf_calib_mfma(int form, int iters, int blocks, double *tflops) {
  return guard([&] {
    if (form < 0 || form > 1 || iters < 1 || blocks < 1) throw std::runtime_error("bad arguments");
    *tflops = hn::mfma_issue_rate(form, iters, blocks);
    return 0;
  });
}

struct hipnegf_server_s {
  std::unique_ptr<hn::Server> srv;
};

int hipnegf_struct_generate_kslab(hipnegf_engine e, hipnegf_struct st, int slot,
                                  unsigned long long seed, double kx, double ky, int with_overlap) {
  return guard([&] {
    hn::gen_kslab(*st->st, slot, seed, kx, ky, with_overlap, e->eng->stream());
    HN_CHECK(hipStreamSynchronize(e->eng->stream()));
    return 0;
  });
}

int hipnegf_struct_get_block(hipnegf_struct st, int slot, int kind, int index, double *host) {
  return guard([&] {
    long long n = 0;
    double2 *d = block_ptr(*st->st, slot, kind, index, &n);
    HN_CHECK(hipMemcpy(host, d, n * sizeof(double2), hipMemcpyDeviceToHost));
    return 0;
  });
}

int hipnegf_server_create(int device, const hipnegf_server_opts *o, hipnegf_server *s) {
  return guard([&] {
    hn::ServeOpts so;
    if (o) {
      if (o->workers) so.workers = o->workers;
      so.policy = o->policy;
      so.max_chunk = o->max_chunk;
      if (o->chunk_flops > 0) so.chunk_flops = o->chunk_flops;
      so.mem_budget = (size_t)std::max(0LL, o->mem_budget);
      so.worker_ws = (size_t)std::max(0LL, o->worker_workspace);
      so.sgf_cache = (size_t)std::max(0LL, o->sgf_cache);
      so.keep_structs = o->keep_structures ? 1 : 0;
    }
    auto *x = new hipnegf_server_s;
    try {
      x->srv.reset(new hn::Server(device, so));
    } catch (...) {
      delete x;
      throw;
    }
    *s = x;
    return 0;
  });
}

int hipnegf_server_destroy(hipnegf_server s) {
  return guard([&] { delete s; return 0; });
}

int hipnegf_server_submit_kslab(hipnegf_server s, unsigned long long seed, int n, int nbl,
                                double kx, double ky, int with_overlap, int nE, const double *E,
                                const hipnegf_opts *o, unsigned long long *id) {
  return guard([&] {
    if (nE <= 0) throw std::runtime_error("nE <= 0");
    hn::ReqStruct rs;
    rs.gen = 1; rs.seed = seed; rs.n = n; rs.nbl = nbl; rs.kx = kx; rs.ky = ky;
    rs.overlap = with_overlap ? 1 : 0;
    *id = s->srv->submit(rs, std::vector<double>(E, E + nE), to_opts(o));
    return 0;
  });
}

int hipnegf_server_submit(hipnegf_server s, hipnegf_struct st, int slot, int nE, const double *E,
                          const hipnegf_opts *o, unsigned long long *id) {
  return guard([&] {
    if (nE <= 0) throw std::runtime_error("nE <= 0");
    hn::ReqStruct rs;
    rs.resident = st->st.get();
    rs.slot = slot;
    *id = s->srv->submit(rs, std::vector<double>(E, E + nE), to_opts(o));
    return 0;
  });
}

int hipnegf_server_poll(hipnegf_server s, unsigned long long id, int *done) {
  return guard([&] { *done = s->srv->done(id) ? 1 : 0; return 0; });
}

int hipnegf_server_wait(hipnegf_server s, unsigned long long id, double *T, int *ncyc, int *status,
                        double *times) {
  return guard([&] {
    hn::ReqResult r = s->srv->wait(id);
    if (times) {
      times[0] = r.t_queue;
      times[1] = r.t_latency;
      times[2] = r.t_gpu;
      times[3] = (double)r.chunks;
    }
    if (!r.error.empty()) throw std::runtime_error(r.error);
    if (T) std::copy(r.T.begin(), r.T.end(), T);
    if (ncyc) std::copy(r.ncyc.begin(), r.ncyc.end(), ncyc);
    if (status) std::copy(r.status.begin(), r.status.end(), status);
    return 0;
  });
}

int hipnegf_server_stats(hipnegf_server s, long long *out) {
  return guard([&] { s->srv->stats(out); return 0; });
}

int hipnegf_gemm_force_variant(int variant) {
  return guard([&] {
    if (variant < -1 || variant > 2) throw std::runtime_error("variant must be -1, 0, 1 or 2");
    hn::gemm_force_variant(variant);
    return 0;
  });
}

int hipnegf_gemm_tune_stats(long long *out) {
  return guard([&] { hn::gemm_tune_stats(out); return 0; });
}

int hipnegf_server_cache_stats(hipnegf_server s, long long *out) {
  return guard([&] { s->srv->cache_stats(out); return 0; });
}

}  // extern "C"
