// HIP kernels of the CG hot path for gfx950 (MI355X).
//
// Every kernel is HBM-bound (SpMV: 2 flop per 12 B in FP64), so the design
// goals are coalesced streams, no re-reads, and reductions that do not cost
// a second launch:
//   * CRS SpMV: CSR-stream.  A block stages the products value[e]*x[index[e]]
//     of a whole row block in LDS with fully coalesced loads, then one thread
//     per row adds its products left to right -- the per-row order of
//     SerialCG::matvecKernelCRS (serial/SerialCG.cpp:87-95).
//   * ELL SpMV: one thread per row over column-major slots j*N + i
//     (Matrix.cpp:339-361), coalesced across the wave, 64-bit slot offsets.
//   * Reductions: wave64 shift-down fold, four wave sums added in LDS, one
//     partial per block, and the last block to finish (elected by a ticket
//     counter) folds the partials and finishes the scalar work (alpha, beta,
//     stop test).  The order is fixed by the geometry in cghip_kernels.h.
//   * SpMV blocks walk a contiguous range of row blocks, and hardware block
//     w (which runs on XCD w % 8) is mapped so that each XCD owns a
//     contiguous slice of rows: x-vector neighbours stay in that XCD's L2.
//
// Floating-point contraction is off (build flag and pragma): a*x+y is a
// rounded multiply then a rounded add, as in the reference's CPU kernels.

#include "cghip_kernels.h"

#pragma clang fp contract(off)

namespace cghip {

namespace {

// ---------------------------------------------------------------------------
// Reductions.

// Lane l accumulates lane l+off for off = 32, 16, ..., 1; lane 0 ends with
// the wave's sum.
__device__ inline double wave_fold(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

// Sum over the 256 threads, returned to every thread: wave sums w0..w3 are
// added left to right.  Ends with a barrier so that `lds` can be reused.
__device__ inline double block_fold(double v, double *lds) {
  v = wave_fold(v);
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = v;
  __syncthreads();
  double r = ((lds[0] + lds[1]) + lds[2]) + lds[3];
  __syncthreads();
  return r;
}

// Stores this block's partial(s) and returns true in every thread of the
// block that finished last.  Partials are written with agent-scope atomics
// and a device fence before the ticket increment, so the last block (which
// fences again) reads every partial whichever XCD's L2 wrote it.
template <int K>
__device__ inline bool publish(const double *bp, Reduce red, int G, int g) {
  __shared__ int s_last;
  if (threadIdx.x == 0) {
    __hip_atomic_store(&red.partial0[g], bp[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (K > 1)
      __hip_atomic_store(&red.partial1[g], bp[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __threadfence();
    unsigned t = atomicAdd(red.ticket, 1u);
    s_last = (t == (unsigned)(G - 1));
  }
  __syncthreads();
  bool last = s_last != 0;
  if (last) __threadfence();
  return last;
}

// Thread t adds partials t, t+256, ... left to right; then the block fold.
__device__ inline double final_fold(const double *partial, int G, double *lds) {
  double s = 0.0;
  for (int i = threadIdx.x; i < G; i += kBlock)
    s += __hip_atomic_load(&partial[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return block_fold(s, lds);
}

__device__ inline void reset_ticket(Reduce red) {
  if (threadIdx.x == 0) __hip_atomic_store(red.ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Hardware block w runs on XCD w % 8.  Give XCD x the contiguous logical
// blocks [base_x, base_x + count_x), count_x = ceil((G - x) / 8).
__device__ inline int xcd_block(int w, int G) {
  int xcd = w & 7, idx = w >> 3, base = 0;
  for (int x = 0; x < xcd; x++) base += (G - x + 7) >> 3;
  return base + idx;
}

// ---------------------------------------------------------------------------
// SpMV.

// CSR-stream over the row blocks [lo, hi) of logical block g.  Returns the
// thread's partial of <x, y> (rows it owns, in order) when WITH_DOT.
template <bool WITH_DOT>
__device__ inline double crs_rows(const Crs<double> &A, const double *__restrict__ x,
                                  double *__restrict__ y, int lo, int hi, double *prod,
                                  int *sptr) {
  const int tid = threadIdx.x;
  double tpart = 0.0;
  for (int rb = lo; rb < hi; rb++) {
    const int r0 = A.rb_start[rb], r1 = A.rb_start[rb + 1];
    const int e0 = A.ptr[r0], e1 = A.ptr[r1];
    const int nrows = r1 - r0;
    for (int t = tid; t <= nrows; t += kBlock) sptr[t] = A.ptr[r0 + t] - e0;
    if (e1 - e0 <= kNnzCap) {
      // Stage the products with coalesced loads, four in flight per thread.
      int e = e0 + tid;
      for (; e + 3 * kBlock < e1; e += 4 * kBlock) {
        int c0 = A.index[e], c1 = A.index[e + kBlock], c2 = A.index[e + 2 * kBlock],
            c3 = A.index[e + 3 * kBlock];
        double v0 = A.value[e], v1 = A.value[e + kBlock], v2 = A.value[e + 2 * kBlock],
               v3 = A.value[e + 3 * kBlock];
        prod[e - e0] = v0 * x[c0];
        prod[e - e0 + kBlock] = v1 * x[c1];
        prod[e - e0 + 2 * kBlock] = v2 * x[c2];
        prod[e - e0 + 3 * kBlock] = v3 * x[c3];
      }
      for (; e < e1; e += kBlock) prod[e - e0] = A.value[e] * x[A.index[e]];
      __syncthreads();
      if (tid < nrows) {
        double s = 0.0;
        const int b = sptr[tid], end = sptr[tid + 1];
        for (int j = b; j < end; j++) s += prod[j];
        y[r0 + tid] = s;
        if (WITH_DOT) tpart += x[r0 + tid] * s;
      }
      __syncthreads();
    } else {
      // One row longer than kNnzCap: chunks through LDS, thread 0 adds in order.
      double s = 0.0;
      for (int c = e0; c < e1; c += kNnzCap) {
        const int ce = min(c + kNnzCap, e1);
        for (int e = c + tid; e < ce; e += kBlock) prod[e - c] = A.value[e] * x[A.index[e]];
        __syncthreads();
        if (tid == 0)
          for (int j = 0; j < ce - c; j++) s += prod[j];
        __syncthreads();
      }
      if (tid == 0) {
        y[r0] = s;
        if (WITH_DOT) tpart += x[r0] * s;
      }
    }
  }
  return tpart;
}

// ELL rows [rb*256, rb*256+256) for rb in [lo, hi): thread t owns row
// rb*256 + t and adds its live slots left to right.
template <bool WITH_DOT>
__device__ inline double ell_rows(const Ell<double> &A, const double *__restrict__ x,
                                  double *__restrict__ y, int lo, int hi) {
  double tpart = 0.0;
  const int64_t n = A.n;
  for (int rb = lo; rb < hi; rb++) {
    const int r = rb * kBlock + (int)threadIdx.x;
    if (r >= A.n) continue;
    const int len = A.length[r];
    double s = 0.0;
    int j = 0;
    for (; j + 3 < len; j += 4) {
      const int64_t k0 = (int64_t)j * n + r;
      int c0 = A.index[k0], c1 = A.index[k0 + n], c2 = A.index[k0 + 2 * n], c3 = A.index[k0 + 3 * n];
      double v0 = A.data[k0], v1 = A.data[k0 + n], v2 = A.data[k0 + 2 * n], v3 = A.data[k0 + 3 * n];
      double p0 = v0 * x[c0], p1 = v1 * x[c1], p2 = v2 * x[c2], p3 = v3 * x[c3];
      s += p0;
      s += p1;
      s += p2;
      s += p3;
    }
    for (; j < len; j++) {
      const int64_t k = (int64_t)j * n + r;
      s += A.data[k] * x[A.index[k]];
    }
    y[r] = s;
    if (WITH_DOT) tpart += x[r] * s;
  }
  return tpart;
}

__global__ __launch_bounds__(kBlock) void k_matvec_crs(Crs<double> A, const double *__restrict__ x,
                                                       double *__restrict__ y) {
  __shared__ double prod[kNnzCap];
  __shared__ int sptr[kRowsMax + 1];
  const int G = gridDim.x, g = xcd_block(blockIdx.x, G);
  const int lo = (int)((int64_t)g * A.nrb / G), hi = (int)((int64_t)(g + 1) * A.nrb / G);
  crs_rows<false>(A, x, y, lo, hi, prod, sptr);
}

__global__ __launch_bounds__(kBlock) void k_matvec_ell(Ell<double> A, const double *__restrict__ x,
                                                       double *__restrict__ y) {
  const int G = gridDim.x, g = xcd_block(blockIdx.x, G);
  const int nrb = (A.n + kBlock - 1) / kBlock;
  const int lo = (int)((int64_t)g * nrb / G), hi = (int)((int64_t)(g + 1) * nrb / G);
  ell_rows<false>(A, x, y, lo, hi);
}

// Last block of the SpMV+<p,q> kernel: alpha = rho / <p,q> (CG.cpp:328-334).
__device__ inline void finish_matvec_dot(double tpart, Reduce red, SolveState *st, History h,
                                         int G, int g, bool skip, double *lds) {
  double bp = block_fold(tpart, lds);
  if (publish<1>(&bp, red, G, g)) {
    double pq = final_fold(red.partial0, G, lds);
    if (threadIdx.x == 0) {
      if (skip) {
        st->done = 1;  // CG_MAX_ITER reached (CG.cpp:323)
      } else {
        const int it = st->iteration;
        const double a = st->rho / pq;
        st->dot_pq = pq;
        st->alpha = a;
        if (h.pq) h.pq[it] = pq;
        if (h.alpha) h.alpha[it] = a;
      }
    }
    reset_ticket(red);
  }
}

__global__ __launch_bounds__(kBlock) void k_pcg_matvec_dot_crs(Crs<double> A, const double *__restrict__ p,
                                                               double *__restrict__ q, Reduce red,
                                                               SolveState *st, History h) {
  __shared__ double prod[kNnzCap];
  __shared__ int sptr[kRowsMax + 1];
  __shared__ double lds[4];
  if (st->done) return;
  const bool skip = st->iteration >= st->max_iter;
  const int G = gridDim.x, g = xcd_block(blockIdx.x, G);
  double tpart = 0.0;
  if (!skip) {
    const int lo = (int)((int64_t)g * A.nrb / G), hi = (int)((int64_t)(g + 1) * A.nrb / G);
    tpart = crs_rows<true>(A, p, q, lo, hi, prod, sptr);
  }
  finish_matvec_dot(tpart, red, st, h, G, g, skip, lds);
}

__global__ __launch_bounds__(kBlock) void k_pcg_matvec_dot_ell(Ell<double> A, const double *__restrict__ p,
                                                               double *__restrict__ q, Reduce red,
                                                               SolveState *st, History h) {
  __shared__ double lds[4];
  if (st->done) return;
  const bool skip = st->iteration >= st->max_iter;
  const int G = gridDim.x, g = xcd_block(blockIdx.x, G);
  double tpart = 0.0;
  if (!skip) {
    const int nrb = (A.n + kBlock - 1) / kBlock;
    const int lo = (int)((int64_t)g * nrb / G), hi = (int)((int64_t)(g + 1) * nrb / G);
    tpart = ell_rows<true>(A, p, q, lo, hi);
  }
  finish_matvec_dot(tpart, red, st, h, G, g, skip, lds);
}

// ---------------------------------------------------------------------------
// Vector kernels (grid-stride; thread gid visits gid, gid + T, ... in order).

__global__ __launch_bounds__(kBlock) void k_axpy(int n, double a, const double *__restrict__ x,
                                                 double *__restrict__ y) {
  const int64_t T = (int64_t)gridDim.x * kBlock;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += T) y[i] = y[i] + a * x[i];
}

__global__ __launch_bounds__(kBlock) void k_xpay(int n, const double *__restrict__ x, double a,
                                                 double *__restrict__ y) {
  const int64_t T = (int64_t)gridDim.x * kBlock;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += T) y[i] = x[i] + a * y[i];
}

__global__ __launch_bounds__(kBlock) void k_jacobi(int n, const double *__restrict__ C,
                                                   const double *__restrict__ x, double *__restrict__ y) {
  const int64_t T = (int64_t)gridDim.x * kBlock;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += T) y[i] = C[i] * x[i];
}

__global__ __launch_bounds__(kBlock) void k_dot_tree(int n, const double *__restrict__ a,
                                                     const double *__restrict__ b, Reduce red,
                                                     double *out) {
  __shared__ double lds[4];
  const int G = gridDim.x, g = blockIdx.x;
  const int64_t T = (int64_t)G * kBlock;
  double s = 0.0;
  for (int64_t i = (int64_t)g * kBlock + threadIdx.x; i < n; i += T) s += a[i] * b[i];
  double bp = block_fold(s, lds);
  if (publish<1>(&bp, red, G, g)) {
    double r = final_fold(red.partial0, G, lds);
    if (threadIdx.x == 0) *out = r;
    reset_ticket(red);
  }
}

// One wave: lanes form 64 products at a time, every lane then adds them in
// index order (the shuffle source is wave-uniform).
__global__ __launch_bounds__(64) void k_dot_serial(int n, const double *__restrict__ a,
                                                   const double *__restrict__ b, double *out) {
  const int lane = threadIdx.x;
  double s = 0.0;
  for (int base = 0; base < n; base += 64) {
    const int i = base + lane;
    const double pr = i < n ? a[i] * b[i] : 0.0;
    const int m = min(64, n - base);
    for (int j = 0; j < m; j++) s += __shfl(pr, j, 64);
  }
  if (lane == 0) *out = s;
}

// r = k + (-1) r (xpay(K, -1, R), CG.cpp:295); p = C o r or p = r
// (CG.cpp:297-303); partials of <r,r> and <p,r> (CG.cpp:305-318).
template <bool JAC>
__global__ __launch_bounds__(kBlock) void k_pcg_init(int n, const double *__restrict__ k,
                                                     double *__restrict__ r, double *__restrict__ p,
                                                     const double *__restrict__ C, Reduce red,
                                                     SolveState *st, History h) {
  __shared__ double lds[4];
  const int G = gridDim.x, g = blockIdx.x;
  const int64_t T = (int64_t)G * kBlock;
  double s_rr = 0.0, s_pr = 0.0;
  for (int64_t i = (int64_t)g * kBlock + threadIdx.x; i < n; i += T) {
    const double ri = k[i] + (-1.0) * r[i];
    r[i] = ri;
    const double pi = JAC ? C[i] * ri : ri;
    p[i] = pi;
    s_rr += ri * ri;
    s_pr += pi * ri;
  }
  double bp[2];
  bp[0] = block_fold(s_rr, lds);
  bp[1] = block_fold(s_pr, lds);
  if (publish<2>(bp, red, G, g)) {
    const double r2 = final_fold(red.partial0, G, lds);
    const double pr = final_fold(red.partial1, G, lds);
    if (threadIdx.x == 0) {
      const double rho = JAC ? pr : r2;
      st->r2 = r2;
      st->nrm2_0 = sqrt(r2);
      st->rho = rho;
      st->iteration = 0;
      st->done = 0;
      if (h.r2) h.r2[0] = r2;
      if (h.rho) h.rho[0] = rho;
    }
    reset_ticket(red);
  }
}

// x += a p; r += (-a) q (CG.cpp:340-342); <r,r>, stop test (CG.cpp:344-355);
// z = C o r, rho = <r,z>, b = rho / rho_old (CG.cpp:357-370).
template <bool JAC>
__global__ __launch_bounds__(kBlock) void k_pcg_update(int n, double *__restrict__ x,
                                                       double *__restrict__ r,
                                                       const double *__restrict__ p,
                                                       const double *__restrict__ q,
                                                       const double *__restrict__ C, Reduce red,
                                                       SolveState *st, History h) {
  __shared__ double lds[4];
  if (st->done) return;
  const double a = st->alpha;
  const double na = -a;
  const int G = gridDim.x, g = blockIdx.x;
  const int64_t T = (int64_t)G * kBlock;
  double s_rr = 0.0, s_rz = 0.0;
  for (int64_t i = (int64_t)g * kBlock + threadIdx.x; i < n; i += T) {
    x[i] = x[i] + a * p[i];
    const double ri = r[i] + na * q[i];
    r[i] = ri;
    const double zi = JAC ? C[i] * ri : ri;
    s_rr += ri * ri;
    s_rz += ri * zi;
  }
  double bp[2];
  bp[0] = block_fold(s_rr, lds);
  bp[1] = block_fold(s_rz, lds);
  if (publish<2>(bp, red, G, g)) {
    const double r2 = final_fold(red.partial0, G, lds);
    const double rz = final_fold(red.partial1, G, lds);
    if (threadIdx.x == 0) {
      const int it = st->iteration;
      st->r2 = r2;
      if (h.r2) h.r2[it + 1] = r2;
      const double res = sqrt(r2) / st->nrm2_0;
      st->residual = res;
      if (res <= st->tol) {
        st->done = 1;
      } else {
        const double rho_old = st->rho;
        const double rho = JAC ? rz : r2;
        const double b = rho / rho_old;
        st->rho_old = rho_old;
        st->rho = rho;
        st->beta = b;
        if (h.rho) h.rho[it + 1] = rho;
        if (h.beta) h.beta[it] = b;
      }
      st->iteration = it + 1;
    }
    reset_ticket(red);
  }
}

// p = z + b p with z = C o r recomputed (CG.cpp:375-381).
template <bool JAC>
__global__ __launch_bounds__(kBlock) void k_pcg_xpay(int n, double *__restrict__ p,
                                                     const double *__restrict__ r,
                                                     const double *__restrict__ C,
                                                     const SolveState *st) {
  if (st->done) return;
  const double b = st->beta;
  const int64_t T = (int64_t)gridDim.x * kBlock;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += T) {
    const double zi = JAC ? C[i] * r[i] : r[i];
    p[i] = zi + b * p[i];
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// Host-side helpers and launches.

int crs_rowblocks(int n, const int *ptr, int *starts) {
  int nb = 0, r = 0;
  while (r < n) {
    const int r0 = r;
    r++;
    while (r < n && r - r0 < kRowsMax && ptr[r + 1] - ptr[r0] <= kNnzCap) r++;
    starts[nb++] = r0;
  }
  starts[nb] = n;
  return nb;
}

static inline dim3 vgrid(int n) { return dim3(vec_grid(n)); }

hipError_t matvec_crs(const Crs<double> &A, const double *x, double *y, hipStream_t s) {
  if (A.n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_matvec_crs, dim3(spmv_grid(A.nrb)), dim3(kBlock), 0, s, A, x, y);
  return hipGetLastError();
}

hipError_t matvec_ell(const Ell<double> &A, const double *x, double *y, hipStream_t s) {
  if (A.n == 0) return hipSuccess;
  const int nrb = (A.n + kBlock - 1) / kBlock;
  hipLaunchKernelGGL(k_matvec_ell, dim3(spmv_grid(nrb)), dim3(kBlock), 0, s, A, x, y);
  return hipGetLastError();
}

hipError_t axpy(int n, double a, const double *x, double *y, hipStream_t s) {
  hipLaunchKernelGGL(k_axpy, vgrid(n), dim3(kBlock), 0, s, n, a, x, y);
  return hipGetLastError();
}

hipError_t xpay(int n, const double *x, double a, double *y, hipStream_t s) {
  hipLaunchKernelGGL(k_xpay, vgrid(n), dim3(kBlock), 0, s, n, x, a, y);
  return hipGetLastError();
}

hipError_t jacobi(int n, const double *C, const double *x, double *y, hipStream_t s) {
  hipLaunchKernelGGL(k_jacobi, vgrid(n), dim3(kBlock), 0, s, n, C, x, y);
  return hipGetLastError();
}

hipError_t dot_tree(int n, const double *a, const double *b, Reduce red, double *out,
                    hipStream_t s) {
  hipLaunchKernelGGL(k_dot_tree, vgrid(n), dim3(kBlock), 0, s, n, a, b, red, out);
  return hipGetLastError();
}

hipError_t dot_serial(int n, const double *a, const double *b, double *out, hipStream_t s) {
  hipLaunchKernelGGL(k_dot_serial, dim3(1), dim3(64), 0, s, n, a, b, out);
  return hipGetLastError();
}

hipError_t pcg_init(int n, const double *k, double *r, double *p, const double *C, Reduce red,
                    SolveState *st, History h, hipStream_t s) {
  if (C)
    hipLaunchKernelGGL(k_pcg_init<true>, vgrid(n), dim3(kBlock), 0, s, n, k, r, p, C, red, st, h);
  else
    hipLaunchKernelGGL(k_pcg_init<false>, vgrid(n), dim3(kBlock), 0, s, n, k, r, p, C, red, st, h);
  return hipGetLastError();
}

hipError_t pcg_matvec_dot_crs(const Crs<double> &A, const double *p, double *q, Reduce red,
                              SolveState *st, History h, hipStream_t s) {
  hipLaunchKernelGGL(k_pcg_matvec_dot_crs, dim3(spmv_grid(A.nrb)), dim3(kBlock), 0, s, A, p, q,
                     red, st, h);
  return hipGetLastError();
}

hipError_t pcg_matvec_dot_ell(const Ell<double> &A, const double *p, double *q, Reduce red,
                              SolveState *st, History h, hipStream_t s) {
  const int nrb = (A.n + kBlock - 1) / kBlock;
  hipLaunchKernelGGL(k_pcg_matvec_dot_ell, dim3(spmv_grid(nrb)), dim3(kBlock), 0, s, A, p, q,
                     red, st, h);
  return hipGetLastError();
}

hipError_t pcg_update(int n, double *x, double *r, const double *p, const double *q,
                      const double *C, Reduce red, SolveState *st, History h, hipStream_t s) {
  if (C)
    hipLaunchKernelGGL(k_pcg_update<true>, vgrid(n), dim3(kBlock), 0, s, n, x, r, p, q, C, red,
                       st, h);
  else
    hipLaunchKernelGGL(k_pcg_update<false>, vgrid(n), dim3(kBlock), 0, s, n, x, r, p, q, C, red,
                       st, h);
  return hipGetLastError();
}

hipError_t pcg_xpay(int n, double *p, const double *r, const double *C, const SolveState *st,
                    hipStream_t s) {
  if (C)
    hipLaunchKernelGGL(k_pcg_xpay<true>, vgrid(n), dim3(kBlock), 0, s, n, p, r, C, st);
  else
    hipLaunchKernelGGL(k_pcg_xpay<false>, vgrid(n), dim3(kBlock), 0, s, n, p, r, C, st);
  return hipGetLastError();
}

}  // namespace cghip
