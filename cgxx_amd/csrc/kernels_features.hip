// Cheap matrix statistics for the learned polynomial preconditioner: one
// pass over the FP64 CRS matrix, one thread per row (grid-stride).
//
// Per row i with diagonal d = a_ii and off-diagonal magnitudes |a_ij|:
//   q0 = log d
//   q1 = sum_{j != i} |a_ij| / d          (Gershgorin radius of D^-1 A)
//   q2 = log(max_{j != i} |a_ij| / d)     (strongest coupling)
//   q3 = log(min_{j != i} |a_ij| / d)     (weakest coupling)
//   q4 = q2 - q3                          (row anisotropy)
// q2..q4 only for rows with off-diagonal entries (explicit zeros skipped).
// Each block writes, per quantity, count, sum, sum of squares, min and max,
// then the row-length histogram (lengths 1..15 and >= 16) and the number of
// rows whose diagonal is missing or not positive.  The host folds the
// blocks in ascending order.

#include "device_common.h"
#include "matrix_stats.h"

#pragma clang fp contract(off)

namespace cghip {
namespace {

using namespace dev;

__device__ inline double block_min(double v, double *lds) {
  for (int off = 32; off >= 1; off >>= 1) v = fmin(v, __shfl_down(v, off, 64));
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = v;
  __syncthreads();
  const double r = fmin(fmin(lds[0], lds[1]), fmin(lds[2], lds[3]));
  __syncthreads();
  return r;
}

__device__ inline double block_max(double v, double *lds) {
  for (int off = 32; off >= 1; off >>= 1) v = fmax(v, __shfl_down(v, off, 64));
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = v;
  __syncthreads();
  const double r = fmax(fmax(lds[0], lds[1]), fmax(lds[2], lds[3]));
  __syncthreads();
  return r;
}

__global__ __launch_bounds__(kBlock) void k_matrix_stats(Crs<double> A, double *__restrict__ out) {
  __shared__ double lds[4];
  // counts in 32-bit integers (exact, half the registers of doubles;
  // converted for the folds, which then add the same whole numbers)
  int cnt[kStatQuantities] = {};
  double sum[kStatQuantities] = {}, sq[kStatQuantities] = {};
  double mn[kStatQuantities], mx[kStatQuantities];
  for (int q = 0; q < kStatQuantities; q++) {
    mn[q] = INFINITY;
    mx[q] = -INFINITY;
  }
  int hist[kStatHistBins] = {};
  int bad = 0;
  const int T = gridDim.x * kBlock;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < A.n; i += T) {
    const int e0 = A.ptr[i], e1 = A.ptr[i + 1];
    double d = 0.0, off = 0.0, smax = 0.0, smin = INFINITY;
    bool has_d = false;
    // Entries 8 at a time, every load in flight before the first is used
    // (clamped to the row's last entry, dropped past it), then taken in
    // stored order through selects: a loop of load then use waited for
    // every entry (593 us at 256^3).
    for (int e = e0; e < e1; e += 8) {
      int c[8];
      double v[8];
#pragma unroll
      for (int u = 0; u < 8; u++) {
        const int k = min(e + u, e1 - 1);
        c[u] = A.index[k];
        v[u] = A.value[k];
      }
#pragma unroll
      for (int u = 0; u < 8; u++) {
        const bool live = e + u < e1;
        const bool diag = live & (c[u] == i);
        const bool offd = live & !diag & (v[u] != 0.0);
        const double a = fabs(v[u]);
        d = diag ? v[u] : d;
        has_d = has_d | diag;
        off = offd ? off + a : off;
        smax = offd ? fmax(smax, a) : smax;
        smin = offd ? fmin(smin, a) : smin;
      }
    }
    const int len = e1 - e0;
    hist[len >= kStatHistBins ? kStatHistBins - 1 : (len < 1 ? 0 : len - 1)] += 1;
    if (!has_d || !(d > 0.0)) {
      bad += 1;
      continue;
    }
    double qv[kStatQuantities];
    bool live[kStatQuantities];
    qv[0] = log(d);
    qv[1] = off / d;
    const bool has_off = smax > 0.0;
    qv[2] = has_off ? log(smax / d) : 0.0;
    qv[3] = has_off ? log(smin / d) : 0.0;
    qv[4] = qv[2] - qv[3];
    live[0] = live[1] = true;
    live[2] = live[3] = live[4] = has_off;
    for (int q = 0; q < kStatQuantities; q++) {
      if (!live[q]) continue;
      cnt[q] += 1;
      sum[q] += qv[q];
      sq[q] += qv[q] * qv[q];
      mn[q] = fmin(mn[q], qv[q]);
      mx[q] = fmax(mx[q], qv[q]);
    }
  }
  double *o = out + (size_t)blockIdx.x * kStatWidth;
  for (int q = 0; q < kStatQuantities; q++) {
    const double c = block_fold((double)cnt[q], lds);
    const double s = block_fold(sum[q], lds);
    const double s2 = block_fold(sq[q], lds);
    const double lo = block_min(mn[q], lds);
    const double hi = block_max(mx[q], lds);
    if (threadIdx.x == 0) {
      o[5 * q + 0] = c;
      o[5 * q + 1] = s;
      o[5 * q + 2] = s2;
      o[5 * q + 3] = lo;
      o[5 * q + 4] = hi;
    }
  }
  for (int b = 0; b < kStatHistBins; b++) {
    const double h = block_fold((double)hist[b], lds);
    if (threadIdx.x == 0) o[5 * kStatQuantities + b] = h;
  }
  const double nb = block_fold((double)bad, lds);
  if (threadIdx.x == 0) o[5 * kStatQuantities + kStatHistBins] = nb;
}

}  // namespace

int matrix_stats_grid(int n) {
  int64_t g = ((int64_t)n + kBlock - 1) / kBlock;
  if (g > kVecGMax) g = kVecGMax;
  if (g < 1) g = 1;
  return (int)g;
}

hipError_t matrix_stats(const Crs<double> &A, double *partials, hipStream_t s) {
  hipLaunchKernelGGL(k_matrix_stats, dim3(matrix_stats_grid(A.n)), dim3(kBlock), 0, s, A, partials);
  return hipGetLastError();
}

}  // namespace cghip
