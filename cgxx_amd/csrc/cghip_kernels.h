// Internal interface between the C ABI (cghip.hip) and the kernels
// (kernels.hip).  Not installed; the public boundary is include/cghip.h.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace cghip {

// Reduction geometry.  The summation order of every reduction is a function
// of these constants and of the vector length / CRS row blocks only -- never
// of the device -- so results are reproducible run to run and restated by
// the CPU oracle.
constexpr int kBlock = 256;      // threads per block (4 wave64s)
constexpr int kVecGMax = 2048;   // max blocks of grid-stride vector kernels
constexpr int kRowsMax = 256;    // max rows in one CRS row block (one per thread)
constexpr int kNnzCap = 2048;    // max nonzeros staged in LDS per row block
constexpr int kSpmvGMax = 2048;  // max blocks of the SpMV kernels

// Vector kernels work on element pairs: G = ceil(pairs / 256), at most 2048.
inline int vec_grid(int n) {
  const int64_t pairs = ((int64_t)n + 1) / 2;
  int64_t g = (pairs + kBlock - 1) / kBlock;
  if (g > kVecGMax) g = kVecGMax;
  if (g < 1) g = 1;
  return (int)g;
}

// SpMV kernels: block g handles row blocks g, g + G, g + 2G, ...
inline int spmv_grid(int nrb) {
  int64_t g = nrb;
  if (g > kSpmvGMax) g = kSpmvGMax;
  if (g < 1) g = 1;
  return (int)g;
}

// Scalars of the fused PCG loop (CG.cpp:288-291 plus the loop counter and
// the stop flag).  No kernel reads a field that the same kernel writes:
// the p-update kernel (last of an iteration) writes `next`, the SpMV kernel
// (first of an iteration) copies `next` to `cur`, and the x/r-update and
// p-update kernels read `cur`.
struct Scalars {
  double rho, r2, residual, beta;
  int iteration, done;
};

struct SolveState {
  Scalars cur, next;
  double nrm2_0, tol;  // written once, before the loop
  int max_iter, pad;
};

// Per-iteration history (nullable arrays of max_iter + 1 entries).
struct History {
  double *r2, *rho, *pq, *alpha, *beta;
};

// Per-block partial sums of the reductions.  A kernel writes its blocks'
// partials with plain stores; the next kernel on the stream (after the
// kernel-boundary coherence point) folds them in every block, in the fixed
// order, so no cross-block handoff ever happens inside a kernel.
struct Partials {
  double *a, *b;  // kVecGMax entries each
};

template <typename V>
struct Crs {
  const int *ptr, *index;
  const V *value;
  const int *rb_start;  // row-block starts, nrb + 1 entries
  int n, nrb;
};

template <typename V>
struct Ell {
  const int *length, *index;
  const V *data;
  int n, width;
};

// Row blocks of the CRS kernels: greedy, at most kRowsMax rows and at most
// kNnzCap nonzeros, except that a block always takes its first row (a row
// longer than kNnzCap is then streamed through LDS in chunks).
int crs_rowblocks(int n, const int *ptr, int *starts);

// ---- launches (all on `s`; each returns the launch status) ----------------
hipError_t matvec_crs(const Crs<double> &A, const double *x, double *y, hipStream_t s);
hipError_t matvec_ell(const Ell<double> &A, const double *x, double *y, hipStream_t s);
hipError_t axpy(int n, double a, const double *x, double *y, hipStream_t s);
hipError_t xpay(int n, const double *x, double a, double *y, hipStream_t s);
hipError_t jacobi(int n, const double *C, const double *x, double *y, hipStream_t s);
hipError_t fill(int n, double v, double *y, hipStream_t s);
hipError_t dot_serial(int n, const double *a, const double *b, double *out, hipStream_t s);
// Tree-ordered dot: partials kernel + one-block fold into out[0].
hipError_t dot_tree(int n, const double *a, const double *b, Partials pa, double *out,
                    hipStream_t s);
// out[0] = sum (b - q)^2, out[1] = sum b^2 (vector-kernel tree order)
hipError_t residual_norms(int n, const double *b, const double *q, Partials pa, double *out,
                          hipStream_t s);

// Fused PCG loop kernels (one iteration = matvec_dot, update, xpay).
hipError_t pcg_init(int n, const double *k, double *r, double *p, const double *C, Partials pa,
                    SolveState *st, History h, hipStream_t s);
hipError_t pcg_matvec_dot_crs(const Crs<double> &A, const double *p, double *q, Partials pq,
                              SolveState *st, hipStream_t s);
hipError_t pcg_matvec_dot_ell(const Ell<double> &A, const double *p, double *q, Partials pq,
                              SolveState *st, hipStream_t s);
hipError_t pcg_update(int n, int gpq, double *x, double *r, const double *p, const double *q,
                      const double *C, Partials pq, Partials pr, SolveState *st, History h,
                      hipStream_t s);
hipError_t pcg_xpay(int n, double *p, const double *r, const double *C, Partials pr,
                    SolveState *st, History h, hipStream_t s);

}  // namespace cghip
