/*
 * cghip.h -- C ABI of the MI355X (gfx950) conjugate-gradient backend.
 *
 * This is the drop-in boundary for the hot path of CGxx: the kernel hooks of
 * the abstract CG class (CG.h:224-237) and the device data set-up of a GPU
 * backend (cuda/CGCUDA.cu:38-112).  The C++ backend class CGHIP
 * (cgxx_amd/host/CGHIP.cpp) maps every hook onto one call below; the Python
 * package binds the same symbols with ctypes.
 *
 * Conventions
 *   - Every function returns an int status: CGHIP_OK (0) or a CGHIP_ERR_*.
 *     cghip_last_error(ctx) returns a message for the last failure.  Nothing
 *     throws across the ABI and nothing exits the process.
 *   - Vectors are named by id (CG::Vector, CG.h:22-35), never by pointer.
 *     The context owns every device buffer; the caller owns host buffers.
 *   - One HIP stream per context.  Every call returns after its device work
 *     is complete (the reference's hooks are synchronous: CGCUDA.cu:139-140).
 *   - A context is not thread-safe; one host thread drives it, as one host
 *     thread drives CG::solve (CG.cpp:284-385).
 *   - Sizes are int (int32 indices, as Matrix.h:16-18, 50-54, 77-86);
 *     ELL slot offsets are computed in 64 bits.
 */
#ifndef CGHIP_H
#define CGHIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CGHIP_OK 0
#define CGHIP_ERR_ARG 1     /* invalid argument or malformed matrix */
#define CGHIP_ERR_HIP 2     /* a HIP runtime call failed */
#define CGHIP_ERR_STATE 3   /* call out of order (e.g. no matrix uploaded) */
#define CGHIP_ERR_MEMORY 4  /* device allocation failed */

/* CG::Vector (CG.h:22-35). */
enum {
  CGHIP_VEC_K = 0,
  CGHIP_VEC_X = 1,
  CGHIP_VEC_P = 2,
  CGHIP_VEC_Q = 3,
  CGHIP_VEC_R = 4,
  CGHIP_VEC_Z = 5
};

/* CG::MatrixFormat (CG.h:38-47).  COO is a host-only format (the reference
 * runs it only in SerialCG, serial/SerialCG.cpp:79-85). */
enum { CGHIP_FORMAT_CRS = 1, CGHIP_FORMAT_ELL = 2 };

/* CG::Preconditioner (CG.h:50-55), plus POLY: a polynomial in D^-1 A set by
 * cghip_set_poly (mixed-precision solver only; no reference counterpart). */
enum { CGHIP_PRECOND_NONE = 0, CGHIP_PRECOND_JACOBI = 1, CGHIP_PRECOND_POLY = 2 };

/* Summation order of the dot-product hook.  TREE: deterministic single-pass
 * wave64 tree (fixed order for a given length).  SERIAL: left to right, the
 * order of SerialCG::vectorDotKernel (serial/SerialCG.cpp:145-155); slow,
 * for bit-exact parity runs. */
enum { CGHIP_DOT_TREE = 0, CGHIP_DOT_SERIAL = 1 };

typedef struct cghip_ctx cghip_ctx;

/* --- context ------------------------------------------------------------ */

/* Replaces CGCUDA::init's device selection (cuda/CGCUDA.cu:29-36). */
int cghip_create(int device, cghip_ctx **out);
int cghip_destroy(cghip_ctx *ctx);
/* Message of the last failed call on ctx (ctx may be NULL: last failure of
 * cghip_create). Never NULL. */
const char *cghip_last_error(const cghip_ctx *ctx);
/* Reduction geometry of the kernels (the summation order is a function of
 * these and of the vector length / CRS row pointer only).  ell_gmax is 0:
 * the FP64 SpMV kernels run one block per row block, uncapped. */
int cghip_geometry(int *block, int *vec_gmax, int *rows_max, int *nnz_cap,
                   int *spmv_gmax, int *ell_gmax, int *dia_max);

/* --- data (CGCUDA::doTransferTo / doTransferFrom, cuda/CGCUDA.cu:38-112) - */

/* Allocates the six N-vectors K, X, P, Q, R, Z (zeroed). */
int cghip_set_size(cghip_ctx *ctx, int n);
/* CRS (MatrixDataCRS, Matrix.h:48-71): ptr[n+1], index[nnz], value[nnz].
 * Validated on the host (monotone ptr, ptr[0] = 0, indices in [0, n)). */
int cghip_upload_crs(cghip_ctx *ctx, int n, int nnz, const int32_t *ptr,
                     const int32_t *index, const double *value);
/* ELL (MatrixDataELL, Matrix.h:74-103), column-major: slot (j, i) at
 * j * n + i for j < width; length[i] <= width live slots per row. */
int cghip_upload_ell(cghip_ctx *ctx, int n, int width, const int32_t *length,
                     const int32_t *index, const double *data);
/* Jacobi::C (Preconditioner.h:12-25): C[i] = 1 / A_ii. */
int cghip_upload_jacobi(cghip_ctx *ctx, const double *C);
/* Chooses which uploaded matrix the hooks and the solver use. */
int cghip_set_format(cghip_ctx *ctx, int format);
int cghip_upload_vector(cghip_ctx *ctx, int id, const double *host);
int cghip_download_vector(cghip_ctx *ctx, int id, double *host);
int cghip_set_dot_mode(cghip_ctx *ctx, int mode);
/* Device address of vector `id` (n doubles on the context's device), as
 * CGCUDABase::Device::getVector maps a CG::Vector to its device array
 * (cuda/CGCUDABase.h:74-91).  Valid until the next cghip_set_size with
 * another n or cghip_destroy; work on it must be ordered with the context's
 * calls by the caller. */
int cghip_vector_pointer(cghip_ctx *ctx, int id, double **out);

/* --- hooks (CG.h:224-237) ------------------------------------------------- */

int cghip_cpy(cghip_ctx *ctx, int dst, int src);               /* CG.h:224 */
int cghip_matvec(cghip_ctx *ctx, int x, int y);                /* CG.h:226 y = A x */
int cghip_axpy(cghip_ctx *ctx, double a, int x, int y);        /* CG.h:228 y = a x + y */
int cghip_xpay(cghip_ctx *ctx, int x, double a, int y);        /* CG.h:230 y = x + a y */
int cghip_vector_dot(cghip_ctx *ctx, int a, int b, double *out); /* CG.h:232 */
int cghip_apply_preconditioner(cghip_ctx *ctx, int x, int y);  /* CG.h:235 y = C o x */

/* Allocation-free dot product.  No counterpart in the reference: for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  *out = <a, b> over n
 * doubles, with the kernels and summation order of cghip_vector_dot in
 * CGHIP_DOT_TREE mode (the same result bit for bit); a, b and out are device
 * pointers on the current device; enqueued on `stream` (a hipStream_t; NULL
 * = the null stream), asynchronously.  Calls no hipMalloc, hipFree or
 * allocating library routine: the per-block partials (at most 2048 doubles,
 * 16 KiB) live in a static device array that each GPU holds until the
 * process ends, looked up per device by cghip_create or by the first call.
 * Every partial a call folds is written by that call, so nothing needs
 * clearing, but calls that overlap on one GPU race on the array.  Errors
 * (n < 0, NULL pointers) are reported by cghip_last_error(NULL). */
int cghip_dot_noalloc(const double *a, const double *b, int n, double *out,
                      void *stream);

/* --- fused solver ----------------------------------------------------------
 * The whole loop of CG::solve (CG.cpp:284-385) on the device: three kernels
 * per iteration (SpMV + <p,q>; r update + <r,r> + <r,z>; x += a p and
 * p = C r + b p).
 * Each kernel writes per-block partial sums; the next kernel folds them (in
 * a fixed order) and computes alpha, beta and the stop test itself, so no
 * host round trip is needed per iteration.  Iterations are replayed from a
 * HIP graph.  Uses X (x0 on entry, solution on return) and K (RHS). */
/* Kernel timing of cghip_solve.  NONE: iterations replayed from a HIP graph,
 * no per-kernel timing.  KERNELS: one iteration at a time, each kernel
 * between two HIP events on the context's stream (adds one host round trip
 * per iteration). */
enum { CGHIP_PROFILE_NONE = 0, CGHIP_PROFILE_KERNELS = 1 };

typedef struct {
  int precond;        /* CGHIP_PRECOND_* */
  int max_iter;       /* CG_MAX_ITER, > 0 */
  double tol;         /* CG_TOLERANCE: stop when sqrt(r2)/sqrt(r2_0) <= tol */
  int batch;          /* iterations per graph launch; 0 = automatic */
  int profile;        /* CGHIP_PROFILE_* */
} cghip_solve_opts;

typedef struct {
  int iterations;     /* loop bodies run, counting the converging one */
  double residual;    /* last sqrt(r2)/sqrt(r2_0) */
  double r2;          /* last <r,r> */
  double wall_s;      /* host wall time of the solve, device synchronised */
  /* profile != NONE: summed kernel time (ms) and counts of the kernels that
   * did work (the SpMV + <p,q>, the r update, the x and p update, the
   * set-up) */
  double ms_matvec, ms_update, ms_xpay, ms_init;
  int n_matvec, n_update, n_xpay, n_init;
} cghip_solve_info;

int cghip_solve(cghip_ctx *ctx, const cghip_solve_opts *opts,
                cghip_solve_info *info);

/* --- mixed precision (no counterpart in the reference, whose arithmetic is
 * FP64 throughout, def.h:6) ------------------------------------------------
 * An FP32 inner CG (FP32 matrix values and vectors, int32 indices, every
 * dot accumulated in FP64) with FP64 reliable updates: whenever the inner
 * recursive residual has dropped by `delta` (or reached `tol`), the FP32
 * correction is added to the FP64 solution X, the true residual K - A X is
 * recomputed in FP64 with the FP64 matrix, and the inner iteration goes on
 * from it with its search direction kept.  The stop test is the FP64 true
 * relative residual ||K - A X|| / ||K - A X0|| <= tol.  CRS only.
 * With the Jacobi (or polynomial) preconditioner the inner iteration runs
 * on the scaled operator D^-1/2 A D^-1/2, D = diag(A), stored in FP32
 * without its unit diagonal: Jacobi-PCG applied as a symmetric scaling
 * (the same iterates in exact arithmetic, no preconditioner vector in the
 * inner loop).  PRECOND_NONE runs on the FP32 copy of A. */

/* Builds the FP32 copy of the CRS values, the scaled operator (when every
 * row has exactly one diagonal entry and it is positive; otherwise only
 * PRECOND_NONE is available to cghip_solve_mixed) and the FP32 work
 * vectors.  Call after cghip_upload_crs; cghip_upload_jacobi is not
 * needed (the scaling comes from the matrix's own diagonal).  The scaled
 * operator is stored in DIA when every row's columns are ascending, its
 * off-diagonal entries lie on at most dia_max (cghip_geometry) diagonals,
 * 4 ndiag N + N <= 8 nnz_offdiag + 4 N and every scaled value is below
 * 65520 in magnitude (finite in FP16); else in packed ELL when
 * W N <= 2 (nnz_offdiag + N), W the longest off-diagonal row; else in CRS.
 * The row sums are the same in every layout (stored order). */
int cghip_prepare_fp32(cghip_ctx *ctx);

/* Layout of the scaled FP32 operator: 0 CRS, 1 ELL, 2 DIA; -1 when there is
 * none (cghip_prepare_fp32 not called, or no positive unit scaling). */
#define CGHIP_LAYOUT_CRS 0
#define CGHIP_LAYOUT_ELL 1
#define CGHIP_LAYOUT_DIA 2
int cghip_fp32_layout(const cghip_ctx *ctx, int *layout);

/* Polynomial preconditioner B = p(D^-1 A) D^-1 with
 * p(l) = sum_{j=0}^{degree} coeffs[j] T_j(2 l / b - 1), the Chebyshev basis
 * on [0, b]; b must bound the spectrum of D^-1 A from above (e.g. its
 * Gershgorin bound).  B is SPD when p > 0 on the spectrum: the call rejects
 * (CGHIP_ERR_ARG) coefficients with p <= 0 anywhere on a 4096-point grid of
 * [0, b].  1 <= degree <= 16.  Applied in scaled form, z = p(A_hat) r with
 * A_hat = D^-1/2 A D^-1/2, by Clenshaw's recurrence: one FP32 SpMV per
 * degree with the recurrence fused into the SpMV kernel.  Needs the scaled
 * operator of cghip_prepare_fp32; dropped by the next matrix upload. */
int cghip_set_poly(cghip_ctx *ctx, int degree, const double *coeffs, double b);

/* Precision of the operator values the polynomial's Clenshaw steps read
 * when the scaled operator is stored in DIA: 16 (default: the FP16 copy of
 * the values, 2 bytes per entry) or 32 (the FP32 values the CG operator
 * reads).  With 16 the preconditioner is the polynomial of the FP16-rounded
 * operator, still SPD when p > 0 on its spectrum (cghip_set_poly then checks
 * p on [-2^-10 b, (1 + 2^-10) b]).  Changing it drops the polynomial: call
 * cghip_set_poly again.  ELL and CRS layouts always read FP32 values. */
int cghip_set_poly_values(cghip_ctx *ctx, int bits);

/* Storage of the values the Clenshaw steps read: CGHIP_POLY_VALUES_FP32
 * (the CG operator's FP32 values: ELL, CRS, or bits 32 above),
 * CGHIP_POLY_VALUES_FP16 (the FP16 row records of the DIA layout) or
 * CGHIP_POLY_VALUES_FP16_SYM (symmetric FP16 storage: only the upper
 * diagonals, one stream each, row i reading its lower values from the
 * rows i - o; chosen by cghip_prepare_fp32 when the DIA layout holds the
 * three mirrored diagonal pairs of a 3D 7-point stencil and every FP16
 * value equals its transpose's bit for bit; the products and their order,
 * hence the results, are those of the records, for half the value bytes);
 * -1 without a scaled operator. */
#define CGHIP_POLY_VALUES_FP32 0
#define CGHIP_POLY_VALUES_FP16 1
#define CGHIP_POLY_VALUES_FP16_SYM 2
int cghip_poly_values_layout(const cghip_ctx *ctx, int *layout);

typedef struct {
  int precond;       /* CGHIP_PRECOND_* */
  int max_iter;      /* inner iterations, > 0 */
  double tol;        /* FP64 true relative residual */
  double delta;      /* reliable-update factor in (0, 1); 0 = 0.01 with the
                        polynomial, 0.1 with Jacobi or none */
  int batch;         /* inner iterations per graph launch; 0 = automatic */
  int profile;       /* CGHIP_PROFILE_* */
} cghip_mixed_opts;

typedef struct {
  int iterations;          /* inner FP32 iterations */
  int updates;             /* reliable updates (FP64 residual recomputations) */
  double true_residual;    /* final FP64 ||K - A X|| / ||K - A X0|| */
  double inner_residual;   /* final FP32 recursive relative residual */
  double wall_s;
  /* profile != NONE: summed kernel time (ms) and launch counts of the SpMV +
   * <p,q>, the r update, the y and p update, the reliable updates and (the
   * polynomial) the Clenshaw SpMV steps */
  double ms_matvec, ms_update, ms_xpay, ms_refresh, ms_poly;
  int n_matvec, n_update, n_xpay, n_refresh, n_poly;
  /* reliable updates that also restarted the search direction (the true
   * residual fell by more than 1e3 since the previous update) */
  int restarts;
} cghip_mixed_info;

int cghip_solve_mixed(cghip_ctx *ctx, const cghip_mixed_opts *opts,
                      cghip_mixed_info *info);

/* Per-iteration scalars of the last cghip_solve: which = 0 r2, 1 rho,
 * 2 dot_pq, 3 alpha, 4 beta.  r2/rho have iterations + 1 entries (index 0 is
 * before the loop), the others iterations.  When the loop stopped by
 * convergence, rho[iterations] and beta[iterations - 1] are NaN: the
 * reference computes them only when it goes on (CG.cpp:357-370).  At most
 * the first 2^20 entries of each array are kept (the device buffer is
 * min(max_iter + 1, 2^20) entries, whatever CG_MAX_ITER is).  Returns the
 * count written. */
int cghip_history(cghip_ctx *ctx, int which, double *out, int cap,
                  int *written);

/* Sets every entry of vector `id` to `value` on the device (resetX,
 * CG.cpp:272-275, without a host transfer). */
int cghip_fill_vector(cghip_ctx *ctx, int id, double value);

/* FP64 true residual of vector x against right-hand side b:
 * *norm_r = ||b - A x||_2 and *norm_b = ||b||_2, with tree-ordered sums.
 * Overwrites vector Q with A x. */
int cghip_true_residual(cghip_ctx *ctx, int x, int b, double *norm_r,
                        double *norm_b);

/* Cheap statistics of the CRS matrix (no reference counterpart; inputs of
 * the learned polynomial preconditioner), one pass over the matrix on the
 * device.  Per row: q0 = log a_ii, q1 = sum_{j!=i} |a_ij| / a_ii, q2 = log of
 * the largest and q3 of the smallest nonzero |a_ij| / a_ii (j != i),
 * q4 = q2 - q3 (q2..q4 over rows with off-diagonal entries).  stats[0..19]:
 * mean, std, min, max of q0..q4 (4 per quantity); stats[20..35]: fractions
 * of rows of length 1..15 and >= 16.  CGHIP_ERR_ARG if a row has no positive
 * diagonal. */
#define CGHIP_MATRIX_STATS 36
int cghip_matrix_stats(cghip_ctx *ctx, double *stats, int cap);

/* Average duration (ms) of `reps` back-to-back launches of the matvec kernel
 * x -> y, timed with HIP events on the context's stream. */
int cghip_time_matvec(cghip_ctx *ctx, int x, int y, int reps, double *avg_ms);

/* Average duration (ms) of `reps` back-to-back launches of one hook's device
 * work on the context's vectors, timed with HIP events on the context's
 * stream (no host synchronisation between launches): MATVEC Q = A P (the
 * selected format), AXPY Q += 1e-3 P, XPAY R = P + 0.5 R, DOT <P, R> (the dot
 * mode's order, partials and final fold), JACOBI Z = C o R, CPY Q = P.  The
 * vectors must hold finite values; Q, R and Z are overwritten.  No reference
 * counterpart: the measurement behind the per-hook bandwidths of the
 * reference's timing categories (CG.h:97-128). */
enum { CGHIP_HOOK_MATVEC = 0, CGHIP_HOOK_AXPY = 1, CGHIP_HOOK_XPAY = 2, CGHIP_HOOK_DOT = 3,
       CGHIP_HOOK_JACOBI = 4, CGHIP_HOOK_CPY = 5 };
int cghip_time_hook(cghip_ctx *ctx, int hook, int reps, double *avg_ms);

/* Average duration (ms) of one Clenshaw step of the polynomial set by
 * cghip_set_poly on the scaled FP32 operator: `reps` applications of the
 * whole recurrence back to back on the context's stream between two HIP
 * events, divided by reps x degree.  Overwrites the FP32 work vectors. */
int cghip_time_poly_step(cghip_ctx *ctx, int reps, double *avg_ms);

/* Algorithmic HBM bytes of one iteration of cghip_solve_mixed with the given
 * preconditioner (CGHIP_PRECOND_NONE, _JACOBI or _POLY with the degree of
 * cghip_set_poly) on the prepared FP32 operator: each array a kernel must
 * read or write counted once (the SpMV's x once).  A cost model for choosing
 * between the preconditioners: the kernels stream at similar rates. */
int cghip_mixed_iteration_bytes(const cghip_ctx *ctx, int precond, double *bytes);

#ifdef __cplusplus
}
#endif
#endif
