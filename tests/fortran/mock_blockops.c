/*
 * Host stand-in for the block operations of include/hipnegf.h (section 1),
 * for CPU tests of the Fortran iso_c_binding layer only: "device" pointers
 * are host allocations and every operation is a plain loop.  Linking
 * tests/fortran/blockops_check.f90 against this file checks, without a GPU,
 * that each Fortran interface passes its arguments (by value / by reference,
 * kinds, order) as the C functions expect.  Not part of the library.
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef double _Complex zc;
typedef float _Complex cc;

int hipnegf_createMat(void **d, size_t bytes) { *d = calloc(bytes ? bytes : 1, 1); return *d ? 0 : 1; }
int hipnegf_deleteMat(void **d) { free(*d); *d = NULL; return 0; }
int hipnegf_mallocAsync(void **d, size_t bytes) { return hipnegf_createMat(d, bytes); }
int hipnegf_freeAsync(void **d) { return hipnegf_deleteMat(d); }
int hipnegf_copyMatH2D(void *h, void *d, size_t bytes) { memcpy(d, h, bytes); return 0; }
int hipnegf_copyMatH2D_async(void *h, void *d, size_t bytes) { memcpy(d, h, bytes); return 0; }
int hipnegf_copyMatD2H(void *h, void *d, size_t bytes) { memcpy(h, d, bytes); return 0; }
int hipnegf_meminfo(size_t *f, size_t *t) { *f = (size_t)1 << 30; *t = (size_t)2 << 30; return 0; }
int hipnegf_deviceSynchronize(void) { return 0; }
int hipnegf_getDeviceCount(int *c) { *c = 1; return 0; }
int hipnegf_getDevice(int *d) { *d = 0; return 0; }
int hipnegf_setDevice(int d) { return d == 0 ? 0 : 1; }
int hipnegf_blasInit(void **h) { *h = malloc(8); return 0; }
int hipnegf_blasFinalize(void *h) { free(h); return 0; }
int hipnegf_solverInit(void **h) { *h = malloc(8); return 0; }
int hipnegf_solverFinalize(void *h) { free(h); return 0; }

#define GEN(T, P, R)                                                                          \
  int hipnegf_##P##multMat(void *h, size_t m, size_t n, size_t k, const void *alpha,           \
                           const void *A_, const void *B_, const void *beta, void *C_,         \
                           int dagger) {                                                       \
    (void)h;                                                                                   \
    const T *A = (const T *)A_, *B = (const T *)B_;                                            \
    T *C = (T *)C_, al = *(const T *)alpha, be = *(const T *)beta;                             \
    if (dagger < 0 || dagger > 2) return -1;                                                   \
    for (size_t j = 0; j < n; ++j)                                                             \
      for (size_t i = 0; i < m; ++i) {                                                         \
        T s = 0;                                                                               \
        for (size_t l = 0; l < k; ++l) {                                                       \
          T a = dagger == 1 ? conj(A[l + i * k]) : A[i + l * m];                               \
          T b = dagger == 2 ? conj(B[j + l * n]) : B[l + j * k];                               \
          s += a * b;                                                                          \
        }                                                                                      \
        C[i + j * m] = al * s + be * C[i + j * m];                                             \
      }                                                                                        \
    return 0;                                                                                  \
  }                                                                                            \
  int hipnegf_##P##inverse(void *hb, void *hs, void *A_, void *X_, size_t n) {                 \
    (void)hb; (void)hs;                                                                        \
    T *M = (T *)malloc(sizeof(T) * n * n), *X = (T *)X_;                                       \
    memcpy(M, A_, sizeof(T) * n * n);                                                          \
    for (size_t t = 0; t < n * n; ++t) X[t] = 0;                                               \
    for (size_t i = 0; i < n; ++i) X[i + i * n] = 1;                                           \
    for (size_t c = 0; c < n; ++c) {                                                           \
      size_t p = c;                                                                            \
      for (size_t r = c + 1; r < n; ++r)                                                       \
        if (cabs(M[r + c * n]) > cabs(M[p + c * n])) p = r;                                    \
      if (M[p + c * n] == 0) { free(M); return 1; }                                            \
      for (size_t j = 0; j < n; ++j) {                                                         \
        T t = M[c + j * n]; M[c + j * n] = M[p + j * n]; M[p + j * n] = t;                     \
        t = X[c + j * n]; X[c + j * n] = X[p + j * n]; X[p + j * n] = t;                       \
      }                                                                                        \
      T piv = M[c + c * n];                                                                    \
      for (size_t j = 0; j < n; ++j) { M[c + j * n] /= piv; X[c + j * n] /= piv; }             \
      for (size_t r = 0; r < n; ++r)                                                           \
        if (r != c) {                                                                          \
          T f = M[r + c * n];                                                                  \
          for (size_t j = 0; j < n; ++j) { M[r + j * n] -= f * M[c + j * n];                   \
                                           X[r + j * n] -= f * X[c + j * n]; }                 \
        }                                                                                      \
    }                                                                                          \
    free(M);                                                                                   \
    return 0;                                                                                  \
  }                                                                                            \
  int hipnegf_##P##matsum(void *h, size_t m, size_t n, const void *alpha, const void *A_,      \
                          const void *beta, const void *B_, void *C_, int dagger) {            \
    (void)h;                                                                                   \
    const T *A = (const T *)A_, *B = (const T *)B_;                                            \
    T *C = (T *)C_, al = *(const T *)alpha, be = *(const T *)beta;                             \
    if ((A_ == C_ && dagger == 1) || (B_ == C_ && dagger == 2)) return -1;                     \
    for (size_t j = 0; j < n; ++j)                                                             \
      for (size_t i = 0; i < m; ++i) {                                                         \
        T a = dagger == 1 ? conj(A[j + i * n]) : A[i + j * m];                                 \
        T b = dagger == 2 ? conj(B[j + i * n]) : B[i + j * m];                                 \
        C[i + j * m] = al * a + be * b;                                                        \
      }                                                                                        \
    return 0;                                                                                  \
  }                                                                                            \
  int hipnegf_##P##initmat(void *A_, size_t n) {                                               \
    T *A = (T *)A_;                                                                            \
    for (size_t j = 0; j < n; ++j)                                                             \
      for (size_t i = 0; i < n; ++i) A[i + j * n] = i == j ? 1 : 0;                            \
    return 0;                                                                                  \
  }                                                                                            \
  R hipnegf_##P##trace(void *h, void *A_, size_t n, void *mask, int present) {                 \
    (void)h;                                                                                   \
    const T *A = (const T *)A_;                                                                \
    const int *mk = (const int *)mask;                                                         \
    double s = 0;                                                                              \
    for (size_t i = 0; i < n; ++i)                                                             \
      if (!present || !mk || mk[i]) s += creal(A[i + i * n]);                                  \
    return (R)s;                                                                               \
  }                                                                                            \
  int hipnegf_##P##matcopy(void *h, const void *A, void *B, size_t nel) {                      \
    (void)h;                                                                                   \
    memcpy(B, A, sizeof(T) * nel);                                                             \
    return 0;                                                                                  \
  }                                                                                            \
  int hipnegf_##P##asum(void *h, void *A_, R *sum, size_t n) {                                 \
    (void)h;                                                                                   \
    const T *A = (const T *)A_;                                                                \
    double s = 0;                                                                              \
    for (size_t i = 0; i < n; ++i) s += fabs(creal(A[i])) + fabs(cimag(A[i]));                \
    *sum = (R)s;                                                                               \
    return 0;                                                                                  \
  }

GEN(zc, Z, double)
GEN(cc, C, float)

/* Sancho-Rubio with the reference GPU stopping rule, host data in T */
#define DEC(T, P, R)                                                                          \
  int hipnegf_##P##decimation(void *hb, void *hs, void *Go_, void *Ao_, void *Bo_, void *Co_, \
                              size_t n, int tf32, int *ncyc, R acc) {                          \
    (void)tf32;                                                                                \
    size_t nn = n * n, b = sizeof(T) * nn;                                                     \
    T *Ao = malloc(b), *Aos = malloc(b), *Bo = malloc(b), *Co = malloc(b), *Go = malloc(b);    \
    T *GC = malloc(b), *C1 = malloc(b), *GB = malloc(b), *B1 = malloc(b), *W = malloc(b);      \
    T one = 1, zero = 0, mone = -1;                                                            \
    memcpy(Ao, Ao_, b); memcpy(Aos, Ao_, b); memcpy(Bo, Bo_, b); memcpy(Co, Co_, b);           \
    int ok = 0, it;                                                                            \
    for (it = 1; it <= 300; ++it) {                                                            \
      hipnegf_##P##inverse(hb, hs, Ao, Go, n);                                                 \
      hipnegf_##P##multMat(hb, n, n, n, &one, Go, Co, &zero, GC, 0);                           \
      hipnegf_##P##multMat(hb, n, n, n, &one, Co, GC, &zero, C1, 0);                           \
      double s = 0;                                                                            \
      for (size_t t = 0; t < nn; ++t) s += fabs(creal(C1[t])) + fabs(cimag(C1[t]));           \
      if (s <= acc) { if (ok) break; ok = 1; } else ok = 0;                                    \
      hipnegf_##P##multMat(hb, n, n, n, &one, Bo, GC, &zero, W, 0);                            \
      for (size_t t = 0; t < nn; ++t) { Aos[t] -= W[t]; Ao[t] -= W[t]; }                       \
      hipnegf_##P##multMat(hb, n, n, n, &one, Go, Bo, &zero, GB, 0);                           \
      hipnegf_##P##multMat(hb, n, n, n, &one, Bo, GB, &zero, B1, 0);                           \
      hipnegf_##P##multMat(hb, n, n, n, &mone, Co, GB, &one, Ao, 0);                           \
      memcpy(Bo, B1, b); memcpy(Co, C1, b);                                                    \
    }                                                                                          \
    *ncyc = it > 300 ? 300 : it;                                                               \
    hipnegf_##P##inverse(hb, hs, Aos, Go_, n);                                                 \
    free(Ao); free(Aos); free(Bo); free(Co); free(Go); free(GC); free(C1); free(GB);           \
    free(B1); free(W);                                                                         \
    return it > 300 ? 2 : 0;                                                                   \
  }

DEC(zc, Z, double)
DEC(cc, C, float)
