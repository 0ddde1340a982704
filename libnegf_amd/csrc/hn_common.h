// hn_common.h -- shared types and helpers of the hipnegf engine (gfx950).
//
// Complex numbers are stored interleaved (re, im) as double2 / float2,
// matrices column-major with leading dimension = rows, exactly the z_DNS /
// c_DNS layout of the reference (mat_def.F90:199-204, 312-317), so host
// buffers coming through the C ABI need no repacking.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <type_traits>
#include <cstdio>
#include <stdexcept>
#include <string>

namespace hn {

typedef double2 zd;
typedef float2 zf;

#define HN_CHECK(cmd)                                                                        \
  do {                                                                                       \
    hipError_t e_ = (cmd);                                                                   \
    if (e_ != hipSuccess) {                                                                  \
      throw std::runtime_error(std::string("HIP error ") + hipGetErrorString(e_) + " at " + \
                               __FILE__ + ":" + std::to_string(__LINE__));                  \
    }                                                                                        \
  } while (0)

__host__ __device__ inline zd zmake(double r, double i) { return make_double2(r, i); }
__host__ __device__ inline zd zadd(zd a, zd b) { return make_double2(a.x + b.x, a.y + b.y); }
__host__ __device__ inline zd zsub(zd a, zd b) { return make_double2(a.x - b.x, a.y - b.y); }
__host__ __device__ inline zd zmul(zd a, zd b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__host__ __device__ inline zd zconj(zd a) { return make_double2(a.x, -a.y); }
__host__ __device__ inline zd zscale(zd a, double s) { return make_double2(a.x * s, a.y * s); }
__host__ __device__ inline double zabs1(zd a) { return fabs(a.x) + fabs(a.y); }
__host__ __device__ inline double zabs(zd a) { return hypot(a.x, a.y); }
__host__ __device__ inline zd zdiv(zd a, zd b) {
  // Smith's algorithm: robust complex division
  if (fabs(b.x) >= fabs(b.y)) {
    double r = b.y / b.x, d = b.x + r * b.y;
    return make_double2((a.x + r * a.y) / d, (a.y - r * a.x) / d);
  } else {
    double r = b.x / b.y, d = b.y + r * b.x;
    return make_double2((a.x * r + a.y) / d, (a.y * r - a.x) / d);
  }
}
__host__ __device__ inline zd zinv1(zd b) { return zdiv(make_double2(1.0, 0.0), b); }

// Operation codes of the batched GEMM (cu_ZmultMat's "dagger" argument,
// gpuroutines.cu:409-422: 0 none, 1 op(A)=A^H, 2 op(B)=B^H), generalised.
enum Op : int { OP_N = 0, OP_T = 1, OP_C = 2 };

// Per-item status word (replaces the reference's process exit on LAPACK
// info != 0, gpuroutines.cu:70-81, and on decimation non-convergence).
enum ItemStatus : int {
  ST_OK = 0,
  ST_SINGULAR = 1,      // zero pivot in a block inverse
  ST_NOT_CONVERGED = 2, // decimation hit 300 iterations
  ST_NONFINITE = 4,
};

// Launch-time limits
constexpr int kWave = 64;

}  // namespace hn
