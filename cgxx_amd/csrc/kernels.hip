// HIP kernels of the CG hot path for gfx950 (MI355X).
//
// Every kernel is HBM-bound (SpMV: 2 flop per 12 B in FP64), so the design
// goals are many independent bytes in flight per CU, coalesced streams, no
// re-reads, and no cross-block synchronisation inside a kernel:
//   * CRS SpMV: block-level CSR-stream.  A block of 256 threads owns a row
//     block (<= 256 rows, <= 2048 nonzeros); it issues all index/value loads
//     of the row block at once (8 rounds of 256 coalesced threads), then all
//     x gathers, stages the products in LDS, and every thread adds its row's
//     products left to right -- the per-row order of
//     SerialCG::matvecKernelCRS (serial/SerialCG.cpp:87-95).  Blocks walk
//     row blocks g, g + G, g + 2G, ... (G <= 2048), so the partial sums of a
//     fused dot stay few.
//   * ELL SpMV: one thread per row over column-major slots j*N + i
//     (Matrix.cpp:339-361), 8 slots in flight per thread, 64-bit slot offsets.
//   * Vector kernels: 16-byte loads of element pairs, 4 pairs in flight per
//     thread, grid-stride over pairs.
//   * Reductions: wave64 shift-down fold, four wave sums added in LDS, one
//     partial per block written with a plain store.  The next kernel on the
//     stream folds the partials (every block, same fixed order) before it
//     uses the sum.  A device-scope fence inside the hot kernels would write
//     back and invalidate the XCD's L2 once per block; measured on MI355X it
//     cost ~90 us per 940 MB update pass, so none is used.
//
// Floating-point contraction is off (build flag and pragma): a*x+y is a
// rounded multiply then a rounded add, as in the reference's CPU kernels.

#include "device_common.h"

#pragma clang fp contract(off)

namespace cghip {

namespace {

using namespace dev;

template <typename V, typename X>
__global__ __launch_bounds__(kBlock) void k_matvec_crs(Crs<V> A, const X *__restrict__ x,
                                                       X *__restrict__ y) {
  __shared__ X prod[kNnzCap];
  for (int rb = blockIdx.x; rb < A.nrb; rb += gridDim.x) crs_rowblock<V, X, false>(A, x, y, rb, prod);
}

template <typename V, typename X, bool O32>
__global__ __launch_bounds__(kBlock) void k_matvec_ell(Ell<V> A, const X *__restrict__ x,
                                                       X *__restrict__ y) {
  const int nrb = (A.n + kBlock - 1) / kBlock;
  for (int rb = blockIdx.x; rb < nrb; rb += gridDim.x) ell_rowblock<V, X, false, O32>(A, x, y, rb);
}

// First kernel of an iteration: copies the loop scalars `next` -> `cur`
// (block 0), stops at CG_MAX_ITER (CG.cpp:323), else q = A p with the
// partials of <p,q> (CG.cpp:325-328).  Returns false if the iteration is off.
__device__ inline bool begin_iteration(SolveState *st) {
  const Scalars s = st->next;
  // Also when already done: the later kernels of this iteration read `cur`.
  const bool stop = s.done || s.iteration >= st->max_iter;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    Scalars c = s;
    if (stop) c.done = 1;
    st->cur = c;
  }
  return !stop;
}

// One block per row block (not a persistent grid: the hardware keeps new
// blocks streaming while others finish, 334 against 376 us at 256^3 for the
// CRS kernel), one partial of <p,q> per block; k_fold_groups then folds
// them in groups of 256 before the r update reads them.  Consecutive row
// blocks on consecutive XCDs (the dispatch order) measured faster than
// giving each XCD a contiguous range of row blocks (382 vs 402 us at
// 256^3; ELL 338 vs 359 us).
__global__ __launch_bounds__(kBlock) void k_pcg_matvec_dot_crs(Crs<double> A, const double *__restrict__ p,
                                                               double *__restrict__ q, Partials pq,
                                                               SolveState *st) {
  __shared__ double prod[kNnzCap];
  __shared__ double lds[4];
  if (!begin_iteration(st)) return;
  const double t = crs_rowblock<double, double, true>(A, p, q, blockIdx.x, prod);
  const double bp = block_fold(t, lds);
  if (threadIdx.x == 0) pq.a[blockIdx.x] = bp;
}

template <bool O32>
__global__ __launch_bounds__(kBlock) void k_pcg_matvec_dot_ell(Ell<double> A, const double *__restrict__ p,
                                                               double *__restrict__ q, Partials pq,
                                                               SolveState *st) {
  __shared__ double lds[4];
  if (!begin_iteration(st)) return;
  const double t = ell_rowblock<double, double, true, O32>(A, p, q, blockIdx.x);
  const double bp = block_fold(t, lds);
  if (threadIdx.x == 0) pq.a[blockIdx.x] = bp;
}

// Group fold: block g folds partials g*256 .. g*256+255 (zero past G) into
// out[g], so that the next kernel folds ceil(G / 256) values instead of G.
__global__ __launch_bounds__(kBlock) void k_fold_groups(const double *__restrict__ in, int G,
                                                        double *__restrict__ out) {
  __shared__ double lds[4];
  const int i = blockIdx.x * kBlock + threadIdx.x;
  const double s = block_fold(i < G ? in[i] : 0.0, lds);
  if (threadIdx.x == 0) out[blockIdx.x] = s;
}

// ---------------------------------------------------------------------------
// Vector kernels.  Element pairs (2q, 2q+1); thread gid of T = 256*G visits
// pairs gid, gid + T, ... in order, element 2q before 2q+1.  The last pair
// of an odd-length vector holds one element.

__global__ __launch_bounds__(kBlock) void k_axpy(int n, double a, const double *__restrict__ x,
                                                 double *__restrict__ y) {
  Pairs P(n);
  for (; P.q + (kUnroll - 1) * P.T < P.full; P.q += kUnroll * P.T) {
    double2 xv[kUnroll], yv[kUnroll];
#pragma unroll
    for (int u = 0; u < kUnroll; u++) {
      xv[u] = v2(x)[P.q + u * P.T];
      yv[u] = v2(y)[P.q + u * P.T];
    }
#pragma unroll
    for (int u = 0; u < kUnroll; u++) {
      yv[u].x = yv[u].x + a * xv[u].x;
      yv[u].y = yv[u].y + a * xv[u].y;
      v2(y)[P.q + u * P.T] = yv[u];
    }
  }
  for (; P.q < P.npair; P.q += P.T) {
    const int64_t e = 2 * P.q;
    y[e] = y[e] + a * x[e];
    if (e + 1 < n) y[e + 1] = y[e + 1] + a * x[e + 1];
  }
}

__global__ __launch_bounds__(kBlock) void k_xpay(int n, const double *__restrict__ x, double a,
                                                 double *__restrict__ y) {
  Pairs P(n);
  for (; P.q + (kUnroll - 1) * P.T < P.full; P.q += kUnroll * P.T) {
    double2 xv[kUnroll], yv[kUnroll];
#pragma unroll
    for (int u = 0; u < kUnroll; u++) {
      xv[u] = v2(x)[P.q + u * P.T];
      yv[u] = v2(y)[P.q + u * P.T];
    }
#pragma unroll
    for (int u = 0; u < kUnroll; u++) {
      yv[u].x = xv[u].x + a * yv[u].x;
      yv[u].y = xv[u].y + a * yv[u].y;
      v2(y)[P.q + u * P.T] = yv[u];
    }
  }
  for (; P.q < P.npair; P.q += P.T) {
    const int64_t e = 2 * P.q;
    y[e] = x[e] + a * y[e];
    if (e + 1 < n) y[e + 1] = x[e + 1] + a * y[e + 1];
  }
}

__global__ __launch_bounds__(kBlock) void k_jacobi(int n, const double *__restrict__ C,
                                                   const double *__restrict__ x, double *__restrict__ y) {
  Pairs P(n);
  for (; P.q + (kUnroll - 1) * P.T < P.full; P.q += kUnroll * P.T) {
    double2 cv[kUnroll], xv[kUnroll];
#pragma unroll
    for (int u = 0; u < kUnroll; u++) {
      cv[u] = v2(C)[P.q + u * P.T];
      xv[u] = v2(x)[P.q + u * P.T];
    }
#pragma unroll
    for (int u = 0; u < kUnroll; u++) {
      double2 yv;
      yv.x = cv[u].x * xv[u].x;
      yv.y = cv[u].y * xv[u].y;
      v2(y)[P.q + u * P.T] = yv;
    }
  }
  for (; P.q < P.npair; P.q += P.T) {
    const int64_t e = 2 * P.q;
    y[e] = C[e] * x[e];
    if (e + 1 < n) y[e + 1] = C[e + 1] * x[e + 1];
  }
}

__global__ __launch_bounds__(kBlock) void k_fill(int n, double v, double *__restrict__ y) {
  const int64_t T = (int64_t)gridDim.x * kBlock;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += T) y[i] = v;
}

__global__ __launch_bounds__(kBlock) void k_dot_partials(int n, const double *__restrict__ a,
                                                         const double *__restrict__ b, Partials pa) {
  __shared__ double lds[4];
  Pairs P(n);
  double s = 0.0;
  for (; P.q + (kUnroll - 1) * P.T < P.full; P.q += kUnroll * P.T) {
    double2 av[kUnroll], bv[kUnroll];
#pragma unroll
    for (int u = 0; u < kUnroll; u++) {
      av[u] = v2(a)[P.q + u * P.T];
      bv[u] = v2(b)[P.q + u * P.T];
    }
#pragma unroll
    for (int u = 0; u < kUnroll; u++) {
      s += av[u].x * bv[u].x;
      s += av[u].y * bv[u].y;
    }
  }
  for (; P.q < P.npair; P.q += P.T) {
    const int64_t e = 2 * P.q;
    s += a[e] * b[e];
    if (e + 1 < n) s += a[e + 1] * b[e + 1];
  }
  const double bp = block_fold(s, lds);
  if (threadIdx.x == 0) pa.a[blockIdx.x] = bp;
}

__global__ __launch_bounds__(kBlock) void k_residual_partials(int n, const double *__restrict__ b,
                                                              const double *__restrict__ q,
                                                              Partials pa) {
  __shared__ double lds[4];
  Pairs P(n);
  double s_rr = 0.0, s_bb = 0.0;
  for (; P.q < P.npair; P.q += P.T) {
    for (int64_t e = 2 * P.q; e < 2 * P.q + 2 && e < n; e++) {
      const double ri = b[e] - q[e];
      s_rr += ri * ri;
      s_bb += b[e] * b[e];
    }
  }
  const double b0 = block_fold(s_rr, lds);
  const double b1 = block_fold(s_bb, lds);
  if (threadIdx.x == 0) {
    pa.a[blockIdx.x] = b0;
    pa.b[blockIdx.x] = b1;
  }
}

// One block: out[0] = fold(pa.a), out[1] = fold(pa.b) (if two_sums).
__global__ __launch_bounds__(kBlock) void k_fold_out(Partials pa, int G, int two_sums, double *out) {
  __shared__ double lds[4];
  const double s0 = fold_partials(pa.a, G, lds);
  const double s1 = two_sums ? fold_partials(pa.b, G, lds) : 0.0;
  if (threadIdx.x == 0) {
    out[0] = s0;
    if (two_sums) out[1] = s1;
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
                                                     const double *__restrict__ C, Partials pa) {
  __shared__ double lds[4];
  Pairs P(n);
  double s_rr = 0.0, s_pr = 0.0;
  for (; P.q < P.npair; P.q += P.T) {
    for (int64_t e = 2 * P.q; e < 2 * P.q + 2 && e < n; e++) {
      const double ri = k[e] + (-1.0) * r[e];
      r[e] = ri;
      const double pi = JAC ? C[e] * ri : ri;
      p[e] = pi;
      s_rr += ri * ri;
      s_pr += pi * ri;
    }
  }
  const double b0 = block_fold(s_rr, lds);
  const double b1 = block_fold(s_pr, lds);
  if (threadIdx.x == 0) {
    pa.a[blockIdx.x] = b0;
    pa.b[blockIdx.x] = b1;
  }
}

// One block: r2, rho and nrm2_0 = sqrt(r2) before the loop (CG.cpp:305-318).
template <bool JAC>
__global__ __launch_bounds__(kBlock) void k_pcg_init_finish(Partials pa, int G, SolveState *st,
                                                            History h) {
  __shared__ double lds[4];
  double r2, pr;
  fold_partials2(pa.a, G, pa.b, G, lds, r2, pr);
  if (threadIdx.x == 0) {
    Scalars s{};
    s.r2 = r2;
    s.rho = JAC ? pr : r2;
    s.residual = 1.0;
    s.iteration = 0;
    s.done = 0;
    st->next = s;
    st->cur = s;
    st->nrm2_0 = sqrt(r2);
    if (h.r2 && h.cap > 0) h.r2[0] = r2;
    if (h.rho && h.cap > 0) h.rho[0] = s.rho;
  }
}

template <bool JAC>
__device__ inline void update_elem(double &rv, double qv, double cv, double na, double &s_rr,
                                   double &s_rz) {
  rv = rv + na * qv;
  const double zv = JAC ? cv * rv : rv;
  s_rr += rv * rv;
  s_rz += rv * zv;
}
// a = rho / <p,q> (CG.cpp:334), kept in st->alpha; r += (-a) q (CG.cpp:342;
// the x += a p of CG.cpp:340 is done by the last kernel, which reads p
// anyway); partials of <r,r> and <r,z> with z = C o r (CG.cpp:344, 363-366).
// gpq > 0: <p,q> folds the gpq group sums in ppq.a (k_fold_groups).
// gpq = -G (G <= 256, small systems, one launch fewer per iteration): the
// block folds the G SpMV partials as k_fold_groups folds its one group,
// and 0.0 + that sum is what folding the single group sum gives (a sum
// with only zeros added, except that -0.0 becomes +0.0), so the bits are
// those of the two-level fold.
template <bool JAC>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(8))) void k_pcg_update(
    int n, int gpq, double *__restrict__ r, const double *__restrict__ q,
    const double *__restrict__ C, Partials ppq, Partials prr, SolveState *st, History h) {
  __shared__ double lds[4];
  const Scalars c = st->cur;
  if (c.done) return;
  const double pq =
      gpq > 0 ? fold_partials(ppq.a, gpq, lds)
              : 0.0 + block_fold((int)threadIdx.x < -gpq ? ppq.a[threadIdx.x] : 0.0, lds);
  const double a = c.rho / pq;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->alpha = a;
    if (h.pq && c.iteration < h.cap) h.pq[c.iteration] = pq;
    if (h.alpha && c.iteration < h.cap) h.alpha[c.iteration] = a;
  }
  const double na = -a;
  Pairs P(n);
  double s_rr = 0.0, s_rz = 0.0;
  for (; P.q + (kUnroll - 1) * P.T < P.full; P.q += kUnroll * P.T) {
    double2 rv[kUnroll], qv[kUnroll], cv[kUnroll];
#pragma unroll
    for (int u = 0; u < kUnroll; u++) {
      const int64_t j = P.q + u * P.T;
      rv[u] = v2(r)[j];
      qv[u] = v2_stream(q, j);
      if (JAC) cv[u] = v2_stream(C, j);
    }
#pragma unroll
    for (int u = 0; u < kUnroll; u++) {
      const int64_t j = P.q + u * P.T;
      update_elem<JAC>(rv[u].x, qv[u].x, JAC ? cv[u].x : 0.0, na, s_rr, s_rz);
      update_elem<JAC>(rv[u].y, qv[u].y, JAC ? cv[u].y : 0.0, na, s_rr, s_rz);
      v2(r)[j] = rv[u];
    }
  }
  for (; P.q < P.npair; P.q += P.T) {
    for (int64_t e = 2 * P.q; e < 2 * P.q + 2 && e < n; e++) {
      double re = r[e];
      update_elem<JAC>(re, q[e], JAC ? C[e] : 0.0, na, s_rr, s_rz);
      r[e] = re;
    }
  }
  const double b0 = block_fold(s_rr, lds);
  const double b1 = block_fold(s_rz, lds);
  if (threadIdx.x == 0) {
    prr.a[blockIdx.x] = b0;
    prr.b[blockIdx.x] = b1;
  }
}
// Last kernel of an iteration: x += a p with the old p (CG.cpp:340); r2 and
// the stop test (CG.cpp:344-355); if not converged rho = <r,z>,
// b = rho / rho_old (CG.cpp:357-370) and p = z + b p with z = C o r
// recomputed (CG.cpp:375-381).  Writes `next`.
template <bool JAC>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(8))) void k_pcg_xpay(
    int n, double *__restrict__ x, double *__restrict__ p, const double *__restrict__ r,
    const double *__restrict__ C, Partials prr, SolveState *st, History h) {
  __shared__ double lds[4];
  const Scalars c = st->cur;
  if (c.done) return;
  const int G = gridDim.x;  // the update kernel ran on the same grid
  double r2, rz;
  fold_partials2(prr.a, G, prr.b, G, lds, r2, rz);
  const double res = sqrt(r2) / st->nrm2_0;
  const bool converged = res <= st->tol;
  const double rho = JAC ? rz : r2;
  const double b = rho / c.rho;
  const double a = st->alpha;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    const int it = c.iteration;
    Scalars s = c;
    s.r2 = r2;
    s.residual = res;
    s.iteration = it + 1;
    if (h.r2 && it + 1 < h.cap) h.r2[it + 1] = r2;
    // The reference computes rho and beta only when the loop goes on
    // (CG.cpp:357-370): after the converging iteration their entries are NaN.
    const double nan = __builtin_nan("");
    if (h.rho && it + 1 < h.cap) h.rho[it + 1] = converged ? nan : rho;
    if (h.beta && it < h.cap) h.beta[it] = converged ? nan : b;
    if (converged) {
      s.done = 1;
    } else {
      s.rho = rho;
      s.beta = b;
    }
    st->next = s;
  }
  // x += a p always; p = z + b p unless the loop stops here.  Four streams
  // per element: unrolled twice to stay within 64 VGPRs.  x, r and C are
  // streamed non-temporally: only p stays cached for the next SpMV.
  constexpr int U = 2;
  Pairs P(n);
  if (converged) {
    for (; P.q < P.npair; P.q += P.T)
      for (int64_t e = 2 * P.q; e < 2 * P.q + 2 && e < n; e++) x[e] = x[e] + a * p[e];
    return;
  }
  for (; P.q + (U - 1) * P.T < P.full; P.q += U * P.T) {
    double2 xv[U], pv[U], rv[U], cv[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int64_t j = P.q + u * P.T;
      xv[u] = v2_stream(x, j);
      pv[u] = v2(p)[j];
      rv[u] = v2_stream(r, j);
      if (JAC) cv[u] = v2_stream(C, j);
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int64_t j = P.q + u * P.T;
      xv[u].x = xv[u].x + a * pv[u].x;
      xv[u].y = xv[u].y + a * pv[u].y;
      const double zx = JAC ? cv[u].x * rv[u].x : rv[u].x;
      const double zy = JAC ? cv[u].y * rv[u].y : rv[u].y;
      pv[u].x = zx + b * pv[u].x;
      pv[u].y = zy + b * pv[u].y;
      v2_store_stream(x, j, xv[u]);
      v2(p)[j] = pv[u];
    }
  }
  for (; P.q < P.npair; P.q += P.T) {
    for (int64_t e = 2 * P.q; e < 2 * P.q + 2 && e < n; e++) {
      const double pe = p[e];
      x[e] = x[e] + a * pe;
      const double ze = JAC ? C[e] * r[e] : r[e];
      p[e] = ze + b * pe;
    }
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

int crs_short_rows(int n, const int *ptr, int nrb) {
  for (int i = 0; i < n; i++)
    if (ptr[i + 1] - ptr[i] > kShortRow) return 0;
  return nrb == (n + kRowsMax - 1) / kRowsMax;
}

static inline dim3 vgrid(int n) { return dim3(vec_grid(n)); }

template <typename V, typename X>
hipError_t matvec_crs_t(const Crs<V> &A, const X *x, X *y, hipStream_t s) {
  if (A.n == 0) return hipSuccess;
  // no reduction: one block per row block (more waves in flight than the
  // persistent grid of the fused kernels)
  hipLaunchKernelGGL((k_matvec_crs<V, X>), dim3(A.nrb > 0 ? A.nrb : 1), dim3(kBlock), 0, s, A, x,
                     y);
  return hipGetLastError();
}

template <typename V, typename X>
hipError_t matvec_ell_t(const Ell<V> &A, const X *x, X *y, hipStream_t s) {
  if (A.n == 0) return hipSuccess;
  const dim3 grid(ell_chunks(A.n) > 0 ? ell_chunks(A.n) : 1);
  if (ell_o32(A))
    hipLaunchKernelGGL((k_matvec_ell<V, X, true>), grid, dim3(kBlock), 0, s, A, x, y);
  else
    hipLaunchKernelGGL((k_matvec_ell<V, X, false>), grid, dim3(kBlock), 0, s, A, x, y);
  return hipGetLastError();
}

hipError_t matvec_crs(const Crs<double> &A, const double *x, double *y, hipStream_t s) {
  return matvec_crs_t<double, double>(A, x, y, s);
}

hipError_t matvec_ell(const Ell<double> &A, const double *x, double *y, hipStream_t s) {
  return matvec_ell_t<double, double>(A, x, y, s);
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

hipError_t fill(int n, double v, double *y, hipStream_t s) {
  hipLaunchKernelGGL(k_fill, vgrid(n), dim3(kBlock), 0, s, n, v, y);
  return hipGetLastError();
}

hipError_t dot_tree(int n, const double *a, const double *b, Partials pa, double *out,
                    hipStream_t s) {
  hipLaunchKernelGGL(k_dot_partials, vgrid(n), dim3(kBlock), 0, s, n, a, b, pa);
  hipLaunchKernelGGL(k_fold_out, dim3(1), dim3(kBlock), 0, s, pa, vec_grid(n), 0, out);
  return hipGetLastError();
}

hipError_t residual_norms(int n, const double *b, const double *q, Partials pa, double *out,
                          hipStream_t s) {
  hipLaunchKernelGGL(k_residual_partials, vgrid(n), dim3(kBlock), 0, s, n, b, q, pa);
  hipLaunchKernelGGL(k_fold_out, dim3(1), dim3(kBlock), 0, s, pa, vec_grid(n), 1, out);
  return hipGetLastError();
}

hipError_t dot_serial(int n, const double *a, const double *b, double *out, hipStream_t s) {
  hipLaunchKernelGGL(k_dot_serial, dim3(1), dim3(64), 0, s, n, a, b, out);
  return hipGetLastError();
}

hipError_t pcg_init(int n, const double *k, double *r, double *p, const double *C, Partials pa,
                    SolveState *st, History h, hipStream_t s) {
  if (C) {
    hipLaunchKernelGGL(k_pcg_init<true>, vgrid(n), dim3(kBlock), 0, s, n, k, r, p, C, pa);
    hipLaunchKernelGGL(k_pcg_init_finish<true>, dim3(1), dim3(kBlock), 0, s, pa, vec_grid(n), st, h);
  } else {
    hipLaunchKernelGGL(k_pcg_init<false>, vgrid(n), dim3(kBlock), 0, s, n, k, r, p, C, pa);
    hipLaunchKernelGGL(k_pcg_init_finish<false>, dim3(1), dim3(kBlock), 0, s, pa, vec_grid(n), st, h);
  }
  return hipGetLastError();
}

// SpMV + <p,q> partials into pq.a (one per row block), then their group
// sums into pq2 (fold_groups(G) entries).
hipError_t pcg_matvec_dot_crs(const Crs<double> &A, const double *p, double *q, Partials pq,
                              double *pq2, SolveState *st, hipStream_t s) {
  const int G = A.nrb > 0 ? A.nrb : 1;
  hipLaunchKernelGGL(k_pcg_matvec_dot_crs, dim3(G), dim3(kBlock), 0, s, A, p, q, pq, st);
  if (pq2)
    hipLaunchKernelGGL(k_fold_groups, dim3(fold_groups(G)), dim3(kBlock), 0, s, pq.a, G, pq2);
  return hipGetLastError();
}

hipError_t pcg_matvec_dot_ell(const Ell<double> &A, const double *p, double *q, Partials pq,
                              double *pq2, SolveState *st, hipStream_t s) {
  const int G = ell_chunks(A.n) > 0 ? ell_chunks(A.n) : 1;
  if (ell_o32(A))
    hipLaunchKernelGGL(k_pcg_matvec_dot_ell<true>, dim3(G), dim3(kBlock), 0, s, A, p, q, pq, st);
  else
    hipLaunchKernelGGL(k_pcg_matvec_dot_ell<false>, dim3(G), dim3(kBlock), 0, s, A, p, q, pq, st);
  if (pq2)
    hipLaunchKernelGGL(k_fold_groups, dim3(fold_groups(G)), dim3(kBlock), 0, s, pq.a, G, pq2);
  return hipGetLastError();
}

hipError_t pcg_update(int n, int gpq, double *r, const double *q, const double *C, Partials pq,
                      Partials pr, SolveState *st, History h, hipStream_t s) {
  if (C)
    hipLaunchKernelGGL(k_pcg_update<true>, vgrid(n), dim3(kBlock), 0, s, n, gpq, r, q, C, pq, pr,
                       st, h);
  else
    hipLaunchKernelGGL(k_pcg_update<false>, vgrid(n), dim3(kBlock), 0, s, n, gpq, r, q, C, pq, pr,
                       st, h);
  return hipGetLastError();
}

hipError_t pcg_xpay(int n, double *x, double *p, const double *r, const double *C, Partials pr,
                    SolveState *st, History h, hipStream_t s) {
  if (C)
    hipLaunchKernelGGL(k_pcg_xpay<true>, vgrid(n), dim3(kBlock), 0, s, n, x, p, r, C, pr, st, h);
  else
    hipLaunchKernelGGL(k_pcg_xpay<false>, vgrid(n), dim3(kBlock), 0, s, n, x, p, r, C, pr, st, h);
  return hipGetLastError();
}

}  // namespace cghip
