// hn_inv.h -- batched complex block inverse (host interface).
#pragma once
#include "hn_common.h"

namespace hn {

constexpr int kInvSuper = 64;  // max super-panel width: deferred rank-64 updates

// Largest n handled by the single-workgroup LDS kernel.
int inv_lds_max_n(int real_bytes);

// Workspace of the blocked path, for `batch` items of order n.
template <typename R>
struct InvWork {
  typedef typename std::conditional<sizeof(R) == 8, double2, float2>::type CT;
  CT *Rb = nullptr;     // n x kInvSuper per item (pivot columns of A, leading dim n)
  int *used = nullptr;  // n per item
  int *piv = nullptr;   // n per item
  int n = 0, batch = 0;
  int prof_items = -1;  // active items (flop counts of masked launches in profiles)
};
size_t inv_work_bytes(int n, int batch, int real_bytes);
// carve the workspace out of a device buffer of inv_work_bytes() bytes
template <typename R> InvWork<R> inv_work_carve(void *buf, int n, int batch);

// Ainv_b = inverse(A_b), b in [0,batch).  A_b is overwritten (used as the
// working matrix).  Partial pivoting.  Items with active[b]==0 are skipped;
// status[b] |= ST_SINGULAR on a zero pivot (never aborts the process).
void inv_z(int n, double2 *A, long long sA, double2 *Ainv, long long sAinv, int batch,
           const int *active, int *status, const InvWork<double> &w, hipStream_t s);
void inv_c(int n, float2 *A, long long sA, float2 *Ainv, long long sAinv, int batch,
           const int *active, int *status, const InvWork<float> &w, hipStream_t s);

}  // namespace hn
