/*
 * CPU oracle for the CGxx conjugate-gradient hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (the HIP library, the
 * cg_hip executable, the Python package) links, loads or calls this code.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use
 * it, and only as the checker or as the timed CPU baseline.
 *
 * It restates, in plain C, the reference's CPU semantics for the path:
 *   - the Matrix Market reader            Matrix.cpp:13-111, 113-121
 *   - COO -> CRS (stable per-row order)   Matrix.cpp:203-228
 *   - COO -> ELL (column-major slots)     Matrix.cpp:339-361
 *   - the Jacobi preconditioner           Preconditioner.cpp:13-23
 *   - the manufactured RHS k = A*1        CG.cpp:256-261
 *   - the serial kernels                  serial/SerialCG.cpp:75-161
 *   - the OpenMP kernels                  openmp/CGOpenMP.cpp:175-265
 *   - the PCG loop                        CG.cpp:284-385
 *   - the solution check                  CG.cpp:387-408
 * plus restatements of the summation orders used by the HIP kernels, so that
 * the GPU path can be checked bit for bit, and the FP32-inner / FP64
 * refinement solver and the polynomial preconditioner that have no reference
 * counterpart.
 *
 * How each function is pinned (see tests/test_oracle.py):
 *   - reader, converters, Jacobi, RHS, serial kernels, PCG loop (order 0):
 *     bit-exact against golden histories produced by the reference's own
 *     serial binary built from its sources (oracle/Makefile, target ref) and
 *     stored in tests/golden/; plus brute force (dense A*x on tiny matrices),
 *     closed forms (A*1 = row sums, stencil spectra) and CG invariants.
 *   - tree orders (orders 1, 2): equal to the serial order on integer-valued
 *     data where every partial sum is exact; bounded by gamma_N sum|a_i b_i|
 *     against a long-double reference otherwise.
 */
#ifndef CG_ORACLE_H
#define CG_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Geometry of the HIP kernels' deterministic reductions (restated). */
#define ORC_BLOCK 256     /* threads per block of every reducing kernel */
#define ORC_VEC_GMAX 2048 /* max blocks of the grid-stride vector kernels */
#define ORC_ROWS_MAX 256  /* max rows of one SpMV row block (one per thread) */
#define ORC_NNZ_CAP 2048  /* max nonzeros staged per CRS row block */
#define ORC_SPMV_GMAX 2048 /* max blocks of the CRS SpMV and FP32 ELL kernels */
#define ORC_ELL_GMAX 0     /* the FP64 SpMV kernels: one block per row block, no cap */
#define ORC_DIA_GMAX 4096  /* max blocks of the FP32 DIA kernels */
#define ORC_DIA_MAX 8      /* max diagonals of the FP32 DIA layout */

/* Matrix Market COO, after the reader's symmetric mirroring. */
typedef struct {
  int N;
  int nz;
  int *I;
  int *J;
  double *V;
  int *nzPerRow;
} orc_coo;

/* Reader.  Returns 0 on success, else one of the reference's failure modes:
 * 1 cannot open, 2 not "matrix coordinate real", 3 not general/symmetric,
 * 4 not square, 5 malformed entry. */
int orc_mm_read(const char *path, orc_coo *out);
void orc_coo_free(orc_coo *c);
int orc_max_nz(const orc_coo *c);

void orc_coo_to_crs(const orc_coo *c, int *ptr, int *index, double *value);
/* ELL: elements = N * maxNz slots, slot(j, i) = j * N + i; padding = 0. */
void orc_coo_to_ell(const orc_coo *c, int maxNz, int *length, int *index,
                    double *data);
/* Returns the number of rows without a diagonal entry (C[i] left NaN). */
int orc_jacobi_init(const orc_coo *c, double *C);
void orc_rhs_ones(const orc_coo *c, double *k);
int orc_check(int N, const double *x, double checkTolerance);

/* Kernels (serial order). */
void orc_spmv_coo(const orc_coo *c, const double *x, double *y);
void orc_spmv_crs(int N, const int *ptr, const int *index, const double *value,
                  const double *x, double *y);
void orc_spmv_ell(int N, const int *length, const int *index,
                  const double *data, const double *x, double *y);
void orc_axpy(int N, double a, const double *x, double *y);
void orc_xpay(int N, const double *x, double a, double *y);
void orc_jacobi_apply(int N, const double *C, const double *x, double *y);
double orc_dot_serial(int N, const double *a, const double *b);

/* Tree orders of the HIP kernels. */
int orc_vec_grid(int N);
int orc_crs_rowblocks(int N, const int *ptr, int *starts /* >= N+1 */);
double orc_dot_tree(int N, const double *a, const double *b);
/* <p,q> in the order of the fused SpMV+dot kernel over the given row blocks
 * (CRS: orc_crs_rowblocks; ELL: 256-row chunks). */
double orc_rowblock_dot_tree(int nrb, const int *starts, const double *p,
                             const double *q);
/* The FP64 fused loop's order: one block per row block, the block partials
 * folded in groups of 256, the group sums folded as block partials. */
double orc_rowblock_dot_tree_groups(int nrb, const int *starts, const double *p,
                                    const double *q);
/* The same with at most gmax blocks (G = min(nrb, gmax)). */
double orc_rowblock_dot_tree_g(int nrb, const int *starts, const double *p,
                               const double *q, int gmax);
/* The same with xcd = 1: when G is a multiple of 8, block g walks the
 * row blocks x R + (g >> 3), stepping by G / 8, below (x + 1) R, with
 * x = g & 7 and R = ceil(nrb / 8) (the library's XCD-contiguous walk:
 * each of the eight XCDs takes one contiguous eighth); otherwise, and with
 * xcd = 0, block g takes g, g + G, ... */
double orc_rowblock_dot_tree_walk(int nrb, const int *starts, const double *p,
                                  const double *q, int gmax, int xcd);

/* PCG (CG.cpp:284-385).
 * order 0: every dot left to right (SerialCG.cpp:145-155);
 * order 1: the fused HIP loop (SpMV+<p,q>, update+<r,r>+<r,z>, Jacobi-xpay);
 * order 2: the HIP hooks one by one, dots in the vector-kernel tree order;
 * order 3: the OpenMP kernels (CGOpenMP.cpp:188-265), nthreads threads.
 * format 0 COO, 1 CRS, 2 ELL; precond 0 none, 1 Jacobi.
 * x holds x0 on entry and the solution on return; k is the RHS.
 * History arrays (nullable) have max_iter + 1 entries:
 *   r2[0], rho[0]: before the loop; r2[i+1], rho[i+1]: after iteration i;
 *   pq[i], alpha[i], beta[i]: iteration i. */
typedef struct {
  int format;
  int precond;
  int order;
  int nthreads;
  int max_iter;
  double tol;
  /* matrix (only the arrays of `format` are read) */
  int N;
  int nz;
  const orc_coo *coo;
  const int *ptr, *index;
  const double *value;
  int ell_width;
  const int *length, *ell_index;
  const double *ell_data;
  const double *C; /* Jacobi */
} orc_problem;

typedef struct {
  int iterations;
  double residual;
} orc_result;

int orc_pcg(const orc_problem *p, const double *k, double *x, orc_result *res,
            double *h_r2, double *h_rho, double *h_pq, double *h_alpha,
            double *h_beta);

/* FP32 vector kernels work on element quads: G = ceil(quads / 256) <= 2048. */
int orc_vec_grid4(int N);

/* The FP32-inner / FP64 reliable-update PCG of the HIP library (no
 * reference counterpart), restated kernel by kernel with the same casts.
 * precond 0: CG on the FP32 copy of A ((float)a_ij, every stored entry).
 * precond 1 (Jacobi) and 2 (polynomial): CG on the FP32 scaled operator
 *   A_hat = D^-1/2 A D^-1/2 with s_i = 1 / sqrt(a_ii) in FP64, off-diagonal
 *   values (float)((a_ij * s_i) * s_j) and the unit diagonal implicit
 *   (y_i = x_i + sum of the off-diagonal products), i.e. Jacobi-PCG applied
 *   as a symmetric scaling; the inner residual is (float)(s_i * r_i) and the
 *   correction x += s_i * y_i.  Polynomial: z = p(A_hat) r_hat (the matrix
 *   p(D^-1 A) D^-1 in scaled form).
 * FP32 row sums in stored order, FP64 dots of exact float products in the
 * kernels' tree orders, alpha and beta in FP64 rounded to FP32 for the
 * updates, x64 and r = k - A x64 in FP64 at every reliable update, stop on
 * the FP64 true relative residual.  The reliable-update trigger compares
 * the inner residual with its value at the first update; an update whose
 * true residual fell by more than 1 / ORC_RESET_RATIO since the previous
 * one also restarts the search direction.  CRS only.
 * x holds x0 on entry and the solution on return.  History (nullable,
 * max_iter + 1 entries) as for orc_pcg, of the inner recursion.
 * Returns 0, -1 (not CRS), -2 (bad polynomial), -3 (a row without a
 * positive diagonal, scaled operator). */
typedef struct {
  int iterations;
  int updates;
  double true_residual;
  double inner_residual;
} orc_mixed_result;

int orc_mixed_pcg(const orc_problem *p, const double *k, double *x,
                  double delta, orc_mixed_result *res, double *h_r2,
                  double *h_rho, double *h_pq, double *h_alpha,
                  double *h_beta);

/* Polynomial preconditioner of the mixed solver (precond 2):
 * p(l) = sum_j c[j] T_j(2 l / b - 1), FP32 coefficients and
 * two_over_b = (float)(2 / b) as the library rounds them.  The rho = <r, z>
 * sums are in the SpMV row-block tree order (z comes out of the last
 * Clenshaw SpMV). */
#define ORC_POLY_MAX 16
/* A reliable update whose true residual fell below this fraction of the
 * previous one restarts the search direction (p = z). */
#define ORC_RESET_RATIO 1e-3
typedef struct {
  int degree;
  float c[ORC_POLY_MAX + 1];
  float two_over_b;
  int fp32_values; /* 1: the Clenshaw steps read the FP32 values even in DIA
                      (cghip_set_poly_values(ctx, 32)) */
} orc_poly;

/* v rounded to binary16 (nearest, ties to even), returned as a float. */
float orc_round_half(float v);

/* z = p(A_hat) r in FP32 on the scaled operator A_hat = D^-1/2 A D^-1/2
 * (built as in orc_mixed_pcg_poly), with the library's expression order;
 * when the library stores A_hat in DIA, the Clenshaw steps use its values
 * rounded to FP16. */
int orc_poly_apply(int N, const int *ptr, const int *index,
                   const double *value, const float *r32, const orc_poly *pl,
                   float *z);

/* y = A_hat x in FP32 on the scaled operator; half = 1: the FP16-rounded
 * values of the Clenshaw steps when the library would store A_hat in DIA
 * (*is_dia = 1), else the FP32 values. */
int orc_op32_apply(int N, const int *ptr, const int *index, const double *value,
                   const float *x, int half, float *y, int *is_dia);

int orc_mixed_pcg_poly(const orc_problem *p, const orc_poly *pl,
                       const double *k, double *x, double delta,
                       orc_mixed_result *res, double *h_r2, double *h_rho,
                       double *h_pq, double *h_alpha, double *h_beta);

#ifdef __cplusplus
}
#endif
#endif
