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
constexpr int kRowsMax = 256;    // max rows in one CRS row block
constexpr int kNnzCap = 2048;    // max nonzeros staged in LDS per row block
constexpr int kSpmvGMax = 2048;  // max blocks of the SpMV kernels

inline int vec_grid(int n) {
  int64_t g = ((int64_t)n + kBlock - 1) / kBlock;
  if (g > kVecGMax) g = kVecGMax;
  if (g < 1) g = 1;
  return (int)g;
}

inline int spmv_grid(int nrb) { return nrb < kSpmvGMax ? (nrb > 0 ? nrb : 1) : kSpmvGMax; }

// Device-resident scalars of the fused PCG loop (CG.cpp:288-291 plus the
// loop counter and the stop flag).
struct SolveState {
  double rho, rho_old, r2, nrm2_0, dot_pq, alpha, beta, residual, tol;
  int iteration, max_iter, done, pad;
};

// Per-iteration history (nullable arrays of max_iter + 1 entries).
struct History {
  double *r2, *rho, *pq, *alpha, *beta;
};

// Scratch of the single-pass reductions: per-block partials and the ticket
// counter that elects the last block.  The last block resets the ticket, so
// no memset is ever needed between launches.
struct Reduce {
  double *partial0, *partial1;
  unsigned *ticket;
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
hipError_t dot_tree(int n, const double *a, const double *b, Reduce red, double *out,
                    hipStream_t s);
hipError_t dot_serial(int n, const double *a, const double *b, double *out, hipStream_t s);

// Fused PCG loop kernels.
hipError_t pcg_init(int n, const double *k, double *r, double *p, const double *C,
                    Reduce red, SolveState *st, History h, hipStream_t s);
hipError_t pcg_matvec_dot_crs(const Crs<double> &A, const double *p, double *q, Reduce red,
                              SolveState *st, History h, hipStream_t s);
hipError_t pcg_matvec_dot_ell(const Ell<double> &A, const double *p, double *q, Reduce red,
                              SolveState *st, History h, hipStream_t s);
hipError_t pcg_update(int n, double *x, double *r, const double *p, const double *q,
                      const double *C, Reduce red, SolveState *st, History h, hipStream_t s);
hipError_t pcg_xpay(int n, double *p, const double *r, const double *C, const SolveState *st,
                    hipStream_t s);

}  // namespace cghip
