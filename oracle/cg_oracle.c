/*
 * CPU oracle for the CGxx conjugate-gradient hot path.
 * TEST INFRASTRUCTURE ONLY -- the header cg_oracle.h says what it restates
 * and how each part is pinned.  Nothing in the product links or calls it.
 */
#include "cg_oracle.h"

#include <ctype.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* Matrix Market reader: Matrix.cpp:13-111.                                  */

static void lower(char *s) {
  for (; *s; s++) *s = (char)tolower((unsigned char)*s);
}

/* getline(): strips the trailing newline, keeps everything else. */
static int read_line(FILE *f, char **buf, size_t *cap) {
  ssize_t n = getline(buf, cap, f);
  if (n < 0) return 0;
  if (n > 0 && (*buf)[n - 1] == '\n') (*buf)[n - 1] = 0;
  return 1;
}

int orc_mm_read(const char *path, orc_coo *out) {
  memset(out, 0, sizeof(*out));
  FILE *f = fopen(path, "r");
  if (!f) return 1;
  char *line = NULL;
  size_t cap = 0;
  char banner[256] = "", mtx[256] = "", crd[256] = "", type[256] = "",
       storage[256] = "";
  int rc = 0;
  if (!read_line(f, &line, &cap)) line = strdup("");
  sscanf(line, "%255s %255s %255s %255s %255s", banner, mtx, crd, type,
         storage);
  lower(mtx);
  lower(crd);
  lower(type);
  lower(storage);
  /* Matrix.cpp:35-40 */
  if (strcmp(banner, "%%MatrixMarket") || strcmp(mtx, "matrix") ||
      strcmp(crd, "coordinate") || strcmp(type, "real")) {
    rc = 2;
    goto done;
  }
  /* Matrix.cpp:42-49 */
  int symmetric = 0;
  if (!strcmp(storage, "symmetric"))
    symmetric = 1;
  else if (strcmp(storage, "general")) {
    rc = 3;
    goto done;
  }
  /* Matrix.cpp:51-54: skip empty and comment lines */
  do {
    if (!read_line(f, &line, &cap)) {
      rc = 5;
      goto done;
    }
  } while (line[0] == 0 || line[0] == '%');
  int M = 0, N = 0, nz = 0;
  if (sscanf(line, "%d %d %d", &M, &N, &nz) != 3) {
    rc = 5;
    goto done;
  }
  if (N != M) { /* Matrix.cpp:62-65 */
    rc = 4;
    goto done;
  }
  if (symmetric) nz = 2 * nz - N; /* Matrix.cpp:67-70 */
  out->N = N;
  out->nz = nz;
  out->I = (int *)malloc(sizeof(int) * (size_t)(nz > 0 ? nz : 1));
  out->J = (int *)malloc(sizeof(int) * (size_t)(nz > 0 ? nz : 1));
  out->V = (double *)malloc(sizeof(double) * (size_t)(nz > 0 ? nz : 1));
  out->nzPerRow = (int *)calloc((size_t)(N > 0 ? N : 1), sizeof(int));
  /* Matrix.cpp:80-104: 1-based -> 0-based, mirror off-diagonals in place */
  for (int i = 0; i < nz; i++) {
    if (!read_line(f, &line, &cap)) {
      rc = 5;
      goto done;
    }
    int a, b;
    double v;
    if (sscanf(line, "%d %d %lf", &a, &b, &v) != 3) {
      rc = 5;
      goto done;
    }
    a--;
    b--;
    if (a < 0 || a >= N || b < 0 || b >= N) {
      rc = 5;
      goto done;
    }
    out->I[i] = a;
    out->J[i] = b;
    out->V[i] = v;
    out->nzPerRow[a]++;
    if (symmetric && a != b) {
      i++;
      if (i >= nz) {
        rc = 5;
        goto done;
      }
      out->I[i] = b;
      out->J[i] = a;
      out->V[i] = v;
      out->nzPerRow[b]++;
    }
  }
done:
  free(line);
  fclose(f);
  if (rc) orc_coo_free(out);
  return rc;
}

void orc_coo_free(orc_coo *c) {
  free(c->I);
  free(c->J);
  free(c->V);
  free(c->nzPerRow);
  memset(c, 0, sizeof(*c));
}

/* Matrix.cpp:113-121 */
int orc_max_nz(const orc_coo *c) {
  int m = 0;
  for (int i = 0; i < c->N; i++)
    if (c->nzPerRow[i] > m) m = c->nzPerRow[i];
  return m;
}

/* Matrix.cpp:203-228: exclusive scan of nzPerRow, then a stable scatter. */
void orc_coo_to_crs(const orc_coo *c, int *ptr, int *index, double *value) {
  int N = c->N;
  int *fill = (int *)malloc(sizeof(int) * (size_t)(N > 0 ? N : 1));
  ptr[0] = 0;
  for (int i = 0; i < N; i++) {
    fill[i] = ptr[i];
    ptr[i + 1] = ptr[i] + c->nzPerRow[i];
  }
  for (int e = 0; e < c->nz; e++) {
    int r = c->I[e];
    index[fill[r]] = c->J[e];
    value[fill[r]] = c->V[e];
    fill[r]++;
  }
  free(fill);
}

/* Matrix.cpp:339-361, with 64-bit slot offsets (slot = j * N + i). */
void orc_coo_to_ell(const orc_coo *c, int maxNz, int *length, int *index,
                    double *data) {
  int N = c->N;
  int64_t slots = (int64_t)N * maxNz;
  memset(index, 0, sizeof(int) * (size_t)slots);
  memset(data, 0, sizeof(double) * (size_t)slots);
  memcpy(length, c->nzPerRow, sizeof(int) * (size_t)N);
  int *fill = (int *)calloc((size_t)(N > 0 ? N : 1), sizeof(int));
  for (int e = 0; e < c->nz; e++) {
    int r = c->I[e];
    int64_t s = (int64_t)fill[r] * N + r;
    index[s] = c->J[e];
    data[s] = c->V[e];
    fill[r]++;
  }
  free(fill);
}

/* Preconditioner.cpp:13-23: C[i] = 1 / A_ii, the last diagonal entry wins. */
int orc_jacobi_init(const orc_coo *c, double *C) {
  int N = c->N;
  for (int i = 0; i < N; i++) C[i] = NAN;
  for (int e = 0; e < c->nz; e++)
    if (c->I[e] == c->J[e]) C[c->I[e]] = 1 / c->V[e];
  int missing = 0;
  for (int i = 0; i < N; i++)
    if (isnan(C[i])) missing++;
  return missing;
}

/* CG.cpp:256-261: k = A * 1 accumulated in COO order. */
void orc_rhs_ones(const orc_coo *c, double *k) {
  memset(k, 0, sizeof(double) * (size_t)c->N);
  for (int e = 0; e < c->nz; e++) k[c->I[e]] += c->V[e];
}

/* CG.cpp:387-408 */
int orc_check(int N, const double *x, double checkTolerance) {
  int errors = 0;
  for (int i = 0; i < N; i++)
    if (fabs(x[i] - 1.0) > checkTolerance) errors++;
  return errors;
}

/* ------------------------------------------------------------------------ */
/* Serial kernels: serial/SerialCG.cpp:79-161.                               */

void orc_spmv_coo(const orc_coo *c, const double *x, double *y) {
  memset(y, 0, sizeof(double) * (size_t)c->N);
  for (int e = 0; e < c->nz; e++) y[c->I[e]] += c->V[e] * x[c->J[e]];
}

void orc_spmv_crs(int N, const int *ptr, const int *index, const double *value,
                  const double *x, double *y) {
  for (int i = 0; i < N; i++) {
    double t = 0;
    for (int j = ptr[i]; j < ptr[i + 1]; j++) t += value[j] * x[index[j]];
    y[i] = t;
  }
}

void orc_spmv_ell(int N, const int *length, const int *index,
                  const double *data, const double *x, double *y) {
  for (int i = 0; i < N; i++) {
    double t = 0;
    for (int j = 0; j < length[i]; j++) {
      int64_t s = (int64_t)j * N + i;
      t += data[s] * x[index[s]];
    }
    y[i] = t;
  }
}

void orc_axpy(int N, double a, const double *x, double *y) {
  for (int i = 0; i < N; i++) y[i] += a * x[i];
}

void orc_xpay(int N, const double *x, double a, double *y) {
  for (int i = 0; i < N; i++) y[i] = x[i] + a * y[i];
}

void orc_jacobi_apply(int N, const double *C, const double *x, double *y) {
  for (int i = 0; i < N; i++) y[i] = C[i] * x[i];
}

double orc_dot_serial(int N, const double *a, const double *b) {
  double r = 0;
  for (int i = 0; i < N; i++) r += a[i] * b[i];
  return r;
}

/* ------------------------------------------------------------------------ */
/* OpenMP kernels: openmp/CGOpenMP.cpp:188-265 (order depends on threads).   */

static void omp_spmv_crs(int N, const int *ptr, const int *index,
                         const double *value, const double *x, double *y,
                         int nt) {
#pragma omp parallel for num_threads(nt)
  for (int i = 0; i < N; i++) {
    double t = 0;
    for (int j = ptr[i]; j < ptr[i + 1]; j++) t += value[j] * x[index[j]];
    y[i] = t;
  }
}

static void omp_spmv_ell(int N, const int *length, const int *index,
                         const double *data, const double *x, double *y,
                         int nt) {
#pragma omp parallel for num_threads(nt)
  for (int i = 0; i < N; i++) {
    double t = 0;
    for (int j = 0; j < length[i]; j++) {
      int64_t s = (int64_t)j * N + i;
      t += data[s] * x[index[s]];
    }
    y[i] = t;
  }
}

static void omp_axpy(int N, double a, const double *x, double *y, int nt) {
#pragma omp parallel for num_threads(nt)
  for (int i = 0; i < N; i++) y[i] += a * x[i];
}

static void omp_xpay(int N, const double *x, double a, double *y, int nt) {
#pragma omp parallel for num_threads(nt)
  for (int i = 0; i < N; i++) y[i] = x[i] + a * y[i];
}

static void omp_jacobi(int N, const double *C, const double *x, double *y,
                       int nt) {
#pragma omp parallel for num_threads(nt)
  for (int i = 0; i < N; i++) y[i] = C[i] * x[i];
}

static double omp_dot(int N, const double *a, const double *b, int nt) {
  double r = 0;
#pragma omp parallel for reduction(+ : r) num_threads(nt)
  for (int i = 0; i < N; i++) r += a[i] * b[i];
  return r;
}

/* ------------------------------------------------------------------------ */
/* Tree orders of the HIP kernels.                                           */
/* A block of 256 threads = 4 wave64s.  Each wave folds its 64 values with   */
/* shift-down steps 32,16,8,4,2,1 (lane l += lane l+off); lane 0 of each     */
/* wave lands in shared memory and thread 0 adds the four wave sums left to  */
/* right.  Block partials are reduced by the last block to finish: thread t  */
/* adds partials t, t+256, ... left to right, then the same block fold.      */

static double wave_fold(double *t) {
  for (int off = 32; off >= 1; off >>= 1)
    for (int l = 0; l < off; l++) t[l] = t[l] + t[l + off];
  return t[0];
}

static double block_fold(const double *v) {
  double t[64], w[4];
  for (int k = 0; k < 4; k++) {
    memcpy(t, v + 64 * k, sizeof(t));
    w[k] = wave_fold(t);
  }
  return ((w[0] + w[1]) + w[2]) + w[3];
}

static double final_fold(const double *partial, int G) {
  double v[ORC_BLOCK];
  for (int t = 0; t < ORC_BLOCK; t++) {
    double s = 0.0;
    for (int i = t; i < G; i += ORC_BLOCK) s += partial[i];
    v[t] = s;
  }
  return block_fold(v);
}

/* Vector kernels work on element pairs (2q, 2q+1): G = ceil(pairs / 256),
 * at most 2048, T = 256 * G threads; thread g*256+t visits pairs
 * q = g*256+t, +T, +2T, ... and adds element 2q before 2q+1. */
int orc_vec_grid(int N) {
  int64_t pairs = ((int64_t)N + 1) / 2;
  int64_t g = (pairs + ORC_BLOCK - 1) / ORC_BLOCK;
  if (g > ORC_VEC_GMAX) g = ORC_VEC_GMAX;
  if (g < 1) g = 1;
  return (int)g;
}

static double vec_tree(int N, const double *a, const double *b,
                       const double *c /* nullable: a*(c*b) */) {
  int G = orc_vec_grid(N);
  int64_t T = (int64_t)G * ORC_BLOCK;
  int64_t pairs = ((int64_t)N + 1) / 2;
  double *partial = (double *)malloc(sizeof(double) * (size_t)G);
  double v[ORC_BLOCK];
  for (int g = 0; g < G; g++) {
    for (int t = 0; t < ORC_BLOCK; t++) {
      double s = 0.0;
      for (int64_t q = (int64_t)g * ORC_BLOCK + t; q < pairs; q += T)
        for (int64_t i = 2 * q; i < 2 * q + 2 && i < N; i++) {
          double bi = c ? c[i] * b[i] : b[i];
          s += a[i] * bi;
        }
      v[t] = s;
    }
    partial[g] = block_fold(v);
  }
  double r = final_fold(partial, G);
  free(partial);
  return r;
}

double orc_dot_tree(int N, const double *a, const double *b) {
  return vec_tree(N, a, b, NULL);
}

/* Row blocks of the CRS SpMV: greedy, at most ORC_ROWS_MAX rows and at most
 * ORC_NNZ_CAP nonzeros, except that a block always takes its first row. */
int orc_crs_rowblocks(int N, const int *ptr, int *starts) {
  int nb = 0, r = 0;
  while (r < N) {
    int r0 = r;
    r++;
    while (r < N && r - r0 < ORC_ROWS_MAX && ptr[r + 1] - ptr[r0] <= ORC_NNZ_CAP)
      r++;
    starts[nb++] = r0;
  }
  starts[nb] = N;
  return nb;
}

/* SpMV geometry: G = min(nrb, 2048) blocks; block g owns row blocks
 * g, g + G, g + 2G, ... and its thread t owns row start + t of each.  The
 * thread adds p[r]*q[r] over its rows in order (each product added to the
 * running sum from 0.0); then the block and final folds. */
double orc_rowblock_dot_tree(int nrb, const int *starts, const double *p,
                             const double *q) {
  return orc_rowblock_dot_tree_g(nrb, starts, p, q, ORC_SPMV_GMAX);
}

double orc_rowblock_dot_tree_g(int nrb, const int *starts, const double *p,
                               const double *q, int gmax) {
  return orc_rowblock_dot_tree_walk(nrb, starts, p, q, gmax, 0);
}

/* kernels_mixed.hip chunk_walk<XCD> */
double orc_rowblock_dot_tree_walk(int nrb, const int *starts, const double *p,
                                  const double *q, int gmax, int xcd) {
  if (nrb <= 0) return 0.0;
  int G = nrb < gmax ? nrb : gmax;
  double *partial = (double *)malloc(sizeof(double) * (size_t)G);
  double v[ORC_BLOCK];
  for (int g = 0; g < G; g++) {
    int rb0 = g, stride = G, rend = nrb;
    if (xcd && G % 8 == 0) {
      const int x = g & 7, R = (nrb + 7) >> 3;
      rb0 = x * R + (g >> 3);
      stride = G >> 3;
      rend = (x + 1) * R < nrb ? (x + 1) * R : nrb;
    }
    for (int t = 0; t < ORC_BLOCK; t++) v[t] = 0.0;
    for (int rb = rb0; rb < rend; rb += stride)
      for (int t = 0; t < ORC_BLOCK; t++) {
        int r = starts[rb] + t;
        v[t] += r < starts[rb + 1] ? p[r] * q[r] : 0.0;
      }
    partial[g] = block_fold(v);
  }
  double r = final_fold(partial, G);
  free(partial);
  return r;
}

/* The FP64 fused loop's <p,q>: one block per row block (thread t owns row
 * starts[g] + t), the block's fold; those partials folded in groups of 256
 * by one block each (zero past the last); the group sums folded like any
 * block partials (final_fold). */
double orc_rowblock_dot_tree_groups(int nrb, const int *starts, const double *p,
                                    const double *q) {
  if (nrb <= 0) return 0.0;
  int G2 = (nrb + ORC_BLOCK - 1) / ORC_BLOCK;
  double *part = (double *)malloc(sizeof(double) * (size_t)nrb);
  double *group = (double *)malloc(sizeof(double) * (size_t)G2);
  double v[ORC_BLOCK];
  for (int g = 0; g < nrb; g++) {
    for (int t = 0; t < ORC_BLOCK; t++) {
      int r = starts[g] + t;
      v[t] = r < starts[g + 1] ? p[r] * q[r] : 0.0;
    }
    part[g] = block_fold(v);
  }
  for (int h = 0; h < G2; h++) {
    for (int t = 0; t < ORC_BLOCK; t++) {
      int i = h * ORC_BLOCK + t;
      v[t] = i < nrb ? part[i] : 0.0;
    }
    group[h] = block_fold(v);
  }
  double r = final_fold(group, G2);
  free(part);
  free(group);
  return r;
}

/* ------------------------------------------------------------------------ */
/* PCG: CG.cpp:284-385.                                                      */

typedef struct {
  const orc_problem *p;
  int *starts; /* row blocks of the SpMV (orders 1) */
  int nrb;
} ctx_t;

static void spmv(const ctx_t *c, const double *x, double *y) {
  const orc_problem *p = c->p;
  int nt = p->nthreads > 0 ? p->nthreads : 1;
  switch (p->format) {
  case 0:
    orc_spmv_coo(p->coo, x, y);
    break;
  case 1:
    if (p->order == 3)
      omp_spmv_crs(p->N, p->ptr, p->index, p->value, x, y, nt);
    else
      orc_spmv_crs(p->N, p->ptr, p->index, p->value, x, y);
    break;
  default:
    if (p->order == 3)
      omp_spmv_ell(p->N, p->length, p->ell_index, p->ell_data, x, y, nt);
    else
      orc_spmv_ell(p->N, p->length, p->ell_index, p->ell_data, x, y);
    break;
  }
}

static double dot(const ctx_t *c, const double *a, const double *b) {
  const orc_problem *p = c->p;
  if (p->order == 0) return orc_dot_serial(p->N, a, b);
  if (p->order == 3)
    return omp_dot(p->N, a, b, p->nthreads > 0 ? p->nthreads : 1);
  return orc_dot_tree(p->N, a, b);
}

static void axpy(const ctx_t *c, double a, const double *x, double *y) {
  if (c->p->order == 3)
    omp_axpy(c->p->N, a, x, y, c->p->nthreads > 0 ? c->p->nthreads : 1);
  else
    orc_axpy(c->p->N, a, x, y);
}

static void xpay(const ctx_t *c, const double *x, double a, double *y) {
  if (c->p->order == 3)
    omp_xpay(c->p->N, x, a, y, c->p->nthreads > 0 ? c->p->nthreads : 1);
  else
    orc_xpay(c->p->N, x, a, y);
}

static void jacobi(const ctx_t *c, const double *x, double *y) {
  if (c->p->order == 3)
    omp_jacobi(c->p->N, c->p->C, x, y, c->p->nthreads > 0 ? c->p->nthreads : 1);
  else
    orc_jacobi_apply(c->p->N, c->p->C, x, y);
}

int orc_pcg(const orc_problem *p, const double *k, double *x, orc_result *res,
            double *h_r2, double *h_rho, double *h_pq, double *h_alpha,
            double *h_beta) {
  int N = p->N;
  size_t bytes = sizeof(double) * (size_t)(N > 0 ? N : 1);
  double *P = (double *)malloc(bytes), *Q = (double *)malloc(bytes),
         *R = (double *)malloc(bytes), *Z = (double *)malloc(bytes);
  ctx_t c = {p, NULL, 0};
  int fused = p->order == 1;
  if (fused) {
    if (p->format == 1) {
      c.starts = (int *)malloc(sizeof(int) * (size_t)(N + 1));
      c.nrb = orc_crs_rowblocks(N, p->ptr, c.starts);
    } else if (p->format == 2) {
      c.nrb = (N + ORC_ROWS_MAX - 1) / ORC_ROWS_MAX;
      c.starts = (int *)malloc(sizeof(int) * (size_t)(c.nrb + 1));
      for (int b = 0; b <= c.nrb; b++)
        c.starts[b] = b * ORC_ROWS_MAX < N ? b * ORC_ROWS_MAX : N;
    } else {
      free(P), free(Q), free(R), free(Z);
      return -1;
    }
  }
  int jac = p->precond == 1;
  double rho, rho_old, r2, nrm2_0, dot_pq, a, b;
  double residual = 0;
  int it;

  /* r(0) = k - A x(0) */
  spmv(&c, x, R);
  xpay(&c, k, -1.0, R);
  if (!jac)
    memcpy(P, R, bytes);
  else
    jacobi(&c, R, P);
  if (fused) {
    /* one pass: <r,r> and <p,r> in the vector-kernel tree */
    r2 = orc_dot_tree(N, R, R);
    nrm2_0 = sqrt(r2);
    rho = jac ? orc_dot_tree(N, P, R) : r2;
  } else {
    r2 = dot(&c, R, R);
    nrm2_0 = sqrt(r2);
    rho = jac ? dot(&c, P, R) : r2;
  }
  if (h_r2) h_r2[0] = r2;
  if (h_rho) h_rho[0] = rho;

  for (it = 0; it < p->max_iter; it++) {
    spmv(&c, P, Q);
    dot_pq = fused ? orc_rowblock_dot_tree_groups(c.nrb, c.starts, P, Q) : dot(&c, P, Q);
    a = rho / dot_pq;
    if (h_pq) h_pq[it] = dot_pq;
    if (h_alpha) h_alpha[it] = a;
    axpy(&c, a, P, x);
    axpy(&c, -a, Q, R);
    r2 = fused ? orc_dot_tree(N, R, R) : dot(&c, R, R);
    if (h_r2) h_r2[it + 1] = r2;
    residual = sqrt(r2) / nrm2_0;
    if (residual <= p->tol) {
      it++;
      break;
    }
    rho_old = rho;
    if (!jac) {
      rho = r2;
    } else {
      jacobi(&c, R, Z);
      rho = fused ? vec_tree(N, R, R, p->C) : dot(&c, R, Z);
    }
    if (h_rho) h_rho[it + 1] = rho;
    b = rho / rho_old;
    if (h_beta) h_beta[it] = b;
    xpay(&c, jac ? Z : R, b, P);
  }
  res->iterations = it;
  res->residual = residual;
  free(P), free(Q), free(R), free(Z);
  free(c.starts);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Mixed precision: FP32 inner PCG with FP64 reliable updates.               */

int orc_vec_grid4(int N) {
  int64_t quads = ((int64_t)N + 3) / 4;
  int64_t g = (quads + ORC_BLOCK - 1) / ORC_BLOCK;
  if (g > ORC_VEC_GMAX) g = ORC_VEC_GMAX;
  if (g < 1) g = 1;
  return (int)g;
}

/* sum_i f[i] over quads order of the FP32 vector kernels, f precomputed. */
static double quad_tree(int N, const double *f) {
  int G = orc_vec_grid4(N);
  int64_t T = (int64_t)G * ORC_BLOCK;
  int64_t quads = ((int64_t)N + 3) / 4;
  double *partial = (double *)malloc(sizeof(double) * (size_t)G);
  double v[ORC_BLOCK];
  for (int g = 0; g < G; g++) {
    for (int t = 0; t < ORC_BLOCK; t++) {
      double s = 0.0;
      for (int64_t q = (int64_t)g * ORC_BLOCK + t; q < quads; q += T)
        for (int64_t i = 4 * q; i < 4 * q + 4 && i < N; i++) s += f[i];
      v[t] = s;
    }
    partial[g] = block_fold(v);
  }
  double r = final_fold(partial, G);
  free(partial);
  return r;
}

typedef struct {
  double rho, r2, residual, r2_ref;
  int iteration, done;
} mp_scalars;

/* The FP32 operator of the inner iteration.
 * scaled = 0: A itself, value32 = (float)a_ij, every stored entry.
 * scaled = 1: A_hat = D^-1/2 A D^-1/2 with its diagonal taken as exactly 1
 *   and not stored: s_i = 1 / sqrt(a_ii) (FP64), off-diagonal values
 *   (float)((a_ij * s_i) * s_j) in stored order, diagonal entries dropped.
 * y = op x: t = 0; t += v * x[c] over the stored entries in order (FP32);
 * y_i = scaled ? x_i + t : t. */
typedef struct {
  int N, unit, nrb, gmax;
  int xcd; /* DIA: the dot kernels walk row blocks XCD-contiguously */
  int *ptr, *index, *starts;
  float *value;
  float *value16; /* DIA: the values rounded to FP16 (Clenshaw steps), else NULL */
  double *s;      /* scaled: s_i, else NULL */
} op32;

/* v rounded to the nearest IEEE binary16 value, ties to even (the
 * conversion of the library's DIA builder): 11 significant bits for
 * |v| >= 2^-14, multiples of 2^-24 below, +-inf from 65520 up.  Division
 * and multiplication by powers of two are exact, and nearbyintf rounds to
 * nearest even in the default rounding mode. */
float orc_round_half(float v) {
  if (v != v) return v;
  float a = fabsf(v);
  if (a >= 65520.0f) return copysignf(INFINITY, v);
  float q;
  if (a < 6.103515625e-05f) { /* 2^-14: subnormal halves */
    q = 5.9604644775390625e-08f; /* 2^-24 */
  } else {
    int e;
    frexpf(a, &e); /* a = m 2^e, m in [0.5, 1) */
    q = ldexpf(1.0f, e - 11);
  }
  return copysignf(nearbyintf(a / q) * q, v);
}

/* The library's choice of the DIA layout for the scaled operator
 * (cghip_prepare_fp32, include/cghip.h): every row's columns ascending, the
 * off-diagonal entries on ndiag <= ORC_DIA_MAX distinct diagonals,
 * ndiag N 4 <= 2^32 and N 4 <= 2^32 (32-bit byte offsets), and
 * 4 ndiag N + N <= 8 nnz_offdiag + 4 N.  Returns 1 for DIA. */
static int dia_chosen(const orc_problem *p, int64_t nnz_nd) {
  int N = p->N, ndiag = 0, off[ORC_DIA_MAX];
  for (int i = 0; i < N; i++)
    for (int e = p->ptr[i]; e < p->ptr[i + 1]; e++) {
      if (e > p->ptr[i] && p->index[e] <= p->index[e - 1]) return 0;
      int o = p->index[e] - i, d = 0;
      if (o == 0) continue;
      while (d < ndiag && off[d] != o) d++;
      if (d < ndiag) continue;
      if (ndiag == ORC_DIA_MAX) return 0;
      off[ndiag++] = o;
    }
  int64_t D = ndiag, n = N;
  return D >= 1 && D * n * 4 <= ((int64_t)1 << 32) && n * 4 <= ((int64_t)1 << 32) &&
         4 * D * n + n <= 8 * nnz_nd + 4 * n;
}

static void op32_free(op32 *o) {
  free(o->ptr), free(o->index), free(o->starts), free(o->value), free(o->value16), free(o->s);
}

/* Returns 0, or -3 if scaled and a row has no positive diagonal entry. */
static int op32_build(const orc_problem *p, int scaled, op32 *o) {
  int N = p->N, nz = p->ptr[N];
  memset(o, 0, sizeof(*o));
  o->N = N;
  o->unit = scaled;
  size_t nb = (size_t)(N > 0 ? N : 1);
  o->ptr = (int *)malloc(sizeof(int) * (nb + 1));
  o->index = (int *)malloc(sizeof(int) * (size_t)(nz > 0 ? nz : 1));
  o->value = (float *)malloc(sizeof(float) * (size_t)(nz > 0 ? nz : 1));
  o->starts = (int *)malloc(sizeof(int) * (nb + 1));
  if (scaled) {
    o->s = (double *)malloc(sizeof(double) * nb);
    for (int i = 0; i < N; i++) {
      double d = 0.0;
      int has = 0;
      for (int e = p->ptr[i]; e < p->ptr[i + 1]; e++)
        if (p->index[e] == i) d = p->value[e], has = 1;
      if (!has || !(d > 0.0)) return -3;
      o->s[i] = 1.0 / sqrt(d);
    }
  }
  int k = 0;
  for (int i = 0; i < N; i++) {
    o->ptr[i] = k;
    for (int e = p->ptr[i]; e < p->ptr[i + 1]; e++) {
      int c = p->index[e];
      if (scaled && c == i) continue;
      o->index[k] = c;
      o->value[k] = scaled ? (float)((p->value[e] * o->s[i]) * o->s[c]) : (float)p->value[e];
      k++;
    }
  }
  o->ptr[N] = k;
  o->nrb = orc_crs_rowblocks(N, o->ptr, o->starts);
  o->gmax = ORC_SPMV_GMAX;
  if (scaled) {
    /* the library stores the scaled operator as DIA (up to 4096 blocks) or
     * else as ELL (up to 2048 blocks) when W * N <= 2 (nnz + N), W the
     * longest row, both over 256-row chunks: same row sums (DIA adds a
     * row's entries by ascending diagonal, i.e. in stored order), the dot
     * tree over uniform 256-row blocks */
    int64_t w = 0;
    for (int i = 0; i < N; i++)
      if (o->ptr[i + 1] - o->ptr[i] > w) w = o->ptr[i + 1] - o->ptr[i];
    int dia = dia_chosen(p, k);
    /* ... and only if every value has a finite FP16 copy */
    for (int e = 0; dia && e < k; e++)
      if (!(fabsf(o->value[e]) < 65520.0f)) dia = 0;
    if (dia || w * (int64_t)N <= 2 * ((int64_t)k + N)) {
      o->nrb = (N + ORC_ROWS_MAX - 1) / ORC_ROWS_MAX;
      for (int b = 0; b <= o->nrb; b++)
        o->starts[b] = b * ORC_ROWS_MAX < N ? b * ORC_ROWS_MAX : N;
      if (dia) {
        o->gmax = ORC_DIA_GMAX;
        o->xcd = 1;
        o->value16 = (float *)malloc(sizeof(float) * (size_t)(k > 0 ? k : 1));
        for (int e = 0; e < k; e++) o->value16[e] = orc_round_half(o->value[e]);
      }
    }
  }
  return 0;
}

static float op32_row(const op32 *o, const float *x, int i) {
  float t = 0.0f;
  for (int e = o->ptr[i]; e < o->ptr[i + 1]; e++) t += o->value[e] * x[o->index[e]];
  return o->unit ? x[i] + t : t;
}

/* The operator of the Clenshaw steps: the FP16-rounded values in DIA. */
static float op32_row_poly(const op32 *o, const float *x, int i) {
  if (!o->value16) return op32_row(o, x, i);
  float t = 0.0f;
  for (int e = o->ptr[i]; e < o->ptr[i + 1]; e++) t += o->value16[e] * x[o->index[e]];
  return o->unit ? x[i] + t : t;
}

/* Clenshaw's recurrence for z = p(op) u in FP32 with the expression order
 * of the HIP kernels' epilogues: B_j lives in w[j % 3]; on entry
 * w[degree % 3] holds B_degree = c_degree u; for j = degree-1 .. 0 one
 * operator application t = op B_{j+1}, then s = (2/b) t - B_{j+1},
 * B_j = (c_j u + 2 s) - B_{j+2} (j > 0) or z = (c_0 u + s) - B_2 (j = 0),
 * B_{j+2} = 0 when j + 2 > degree.  z ends in w[0]. */
static void poly_clenshaw(const op32 *o, const float *u, const orc_poly *pl,
                          float *w[3]) {
  const float tb = pl->two_over_b;
  for (int j = pl->degree - 1; j >= 0; j--) {
    const float *b1 = w[(j + 1) % 3];
    const float *b2 = (j + 2 <= pl->degree) ? w[(j + 2) % 3] : NULL;
    float *b0 = w[j % 3];
    const float cj = pl->c[j];
    for (int i = 0; i < o->N; i++) {
      const float t = pl->fp32_values ? op32_row(o, b1, i) : op32_row_poly(o, b1, i);
      const float sv = tb * t - b1[i];
      const float bb = b2 ? b2[i] : 0.0f;
      b0[i] = j > 0 ? (cj * u[i] + 2.0f * sv) - bb : (cj * u[i] + sv) - bb;
    }
  }
}

int orc_poly_apply(int N, const int *ptr, const int *index,
                   const double *value, const float *r32, const orc_poly *pl,
                   float *z) {
  if (!pl || pl->degree < 1 || pl->degree > ORC_POLY_MAX) return -2;
  orc_problem p;
  memset(&p, 0, sizeof(p));
  p.N = N;
  p.ptr = ptr, p.index = index, p.value = value;
  op32 o;
  int rc = op32_build(&p, 1, &o);
  if (rc) {
    op32_free(&o);
    return rc;
  }
  size_t nb = (size_t)(N > 0 ? N : 1);
  float *w[3];
  for (int j = 0; j < 3; j++) w[j] = (float *)calloc(nb, sizeof(float));
  const float ck = pl->c[pl->degree];
  for (int i = 0; i < N; i++) w[pl->degree % 3][i] = ck * r32[i];
  poly_clenshaw(&o, r32, pl, w);
  memcpy(z, w[0], sizeof(float) * (size_t)N);
  for (int j = 0; j < 3; j++) free(w[j]);
  op32_free(&o);
  return 0;
}

/* y = A_hat x in FP32 on the scaled operator (built as in orc_mixed_pcg_poly,
 * implicit unit diagonal): the CG operator with half = 0 (FP32 values), the
 * Clenshaw steps' operator with half = 1 (the FP16-rounded values when the
 * library stores A_hat in DIA; returns 1 in *is_dia then, else 0 and the
 * FP32 values are used).  Returns 0, or -3 without a positive diagonal. */
int orc_op32_apply(int N, const int *ptr, const int *index, const double *value,
                   const float *x, int half, float *y, int *is_dia) {
  orc_problem p;
  memset(&p, 0, sizeof(p));
  p.N = N;
  p.ptr = ptr, p.index = index, p.value = value;
  op32 o;
  int rc = op32_build(&p, 1, &o);
  if (!rc) {
    for (int i = 0; i < N; i++) y[i] = half ? op32_row_poly(&o, x, i) : op32_row(&o, x, i);
    if (is_dia) *is_dia = o.value16 != NULL;
  }
  op32_free(&o);
  return rc;
}

int orc_mixed_pcg(const orc_problem *p, const double *k, double *x,
                  double delta, orc_mixed_result *res, double *h_r2,
                  double *h_rho, double *h_pq, double *h_alpha,
                  double *h_beta) {
  return orc_mixed_pcg_poly(p, NULL, k, x, delta, res, h_r2, h_rho, h_pq,
                            h_alpha, h_beta);
}

int orc_mixed_pcg_poly(const orc_problem *p, const orc_poly *pl,
                       const double *k, double *x, double delta,
                       orc_mixed_result *res, double *h_r2, double *h_rho,
                       double *h_pq, double *h_alpha, double *h_beta) {
  if (p->format != 1) return -1;
  if (p->precond == 2 && (!pl || pl->degree < 1 || pl->degree > ORC_POLY_MAX))
    return -2;
  int N = p->N;
  int poly = p->precond == 2;
  int scaled = p->precond != 0;
  op32 o;
  int rc = op32_build(p, scaled, &o);
  if (rc) {
    op32_free(&o);
    return rc;
  }
  size_t nb = (size_t)(N > 0 ? N : 1);
  float *y32 = (float *)calloc(nb, sizeof(float)), *r32 = (float *)malloc(sizeof(float) * nb),
        *p32 = (float *)malloc(sizeof(float) * nb), *q32 = (float *)malloc(sizeof(float) * nb);
  double *q64 = (double *)malloc(sizeof(double) * nb), *r64 = (double *)malloc(sizeof(double) * nb),
         *f0 = (double *)malloc(sizeof(double) * nb), *f1 = (double *)malloc(sizeof(double) * nb);
  float *w[3] = {NULL, NULL, NULL};
  if (poly)
    for (int j = 0; j < 3; j++) w[j] = (float *)calloc(nb, sizeof(float));
  const int deg = poly ? pl->degree : 0;
  const float ck = poly ? pl->c[deg] : 0.0f;
  /* cghip_solve_mixed's default factor: 0.01 with the polynomial, else 0.1 */
  const double dlt = delta > 0 ? delta : (poly ? 0.01 : 0.1);
  double delta2 = dlt * dlt;
  double nrm2_0 = 0, inner0 = 0, true_res = 0;
  int updates = 0;
  mp_scalars s = {0, 0, 0, 0, 0, 0};

  /* Reliable update: x64 += s o y32 (y32 unscaled when not scaled) unless
   * first; r64 = k - A x64 (FP64); r32 = (float)(s o r64); y32 = 0. */
#define REFRESH(first)                                                        \
  do {                                                                        \
    if (!(first))                                                             \
      for (int i = 0; i < N; i++)                                             \
        x[i] = x[i] + (scaled ? o.s[i] * (double)y32[i] : (double)y32[i]);    \
    orc_spmv_crs(N, p->ptr, p->index, p->value, x, q64);                      \
    for (int i = 0; i < N; i++) r64[i] = k[i] - q64[i];                       \
    double rt2 = orc_dot_tree(N, r64, r64);                                   \
    for (int i = 0; i < N; i++) {                                             \
      float r = scaled ? (float)(o.s[i] * r64[i]) : (float)r64[i];            \
      r32[i] = r;                                                             \
      y32[i] = 0.0f;                                                          \
      if (poly) w[deg % 3][i] = ck * r;                                       \
      else if (first) p32[i] = r;                                             \
      f0[i] = (double)r * (double)r;                                          \
    }                                                                         \
    double rr = quad_tree(N, f0), rz = rr;                                    \
    if (poly) {                                                               \
      poly_clenshaw(&o, r32, pl, w);                                          \
      for (int i = 0; i < N; i++) {                                           \
        if (first) p32[i] = w[0][i];                                          \
        f0[i] = (double)r32[i];                                               \
        f1[i] = (double)w[0][i];                                              \
      }                                                                       \
      rz = orc_rowblock_dot_tree_walk(o.nrb, o.starts, f0, f1, o.gmax, o.xcd);                    \
    }                                                                         \
    if (first) {                                                              \
      nrm2_0 = sqrt(rt2);                                                     \
      inner0 = sqrt(rr);                                                      \
      updates = 0;                                                            \
      s.iteration = 0;                                                        \
    } else {                                                                  \
      updates++;                                                              \
    }                                                                         \
    true_res = sqrt(rt2) / nrm2_0;                                            \
    s.rho = rz;                                                               \
    s.r2 = rr;                                                                \
    s.r2_ref = rr;                                                            \
    s.residual = sqrt(rr) / inner0;                                           \
    s.done = true_res <= p->tol ? 1 : 0;                                      \
    if ((first) && h_r2) h_r2[0] = rr;                                        \
    if ((first) && h_rho) h_rho[0] = rz;                                      \
  } while (0)

  REFRESH(1);
  for (;;) {
    if (s.done == 1) break;
    if (s.done == 2) {
      /* A phase that reduced the true residual by more than 1e3 (e.g. one
       * step that reached FP32 resolution) leaves a search direction that
       * carries no accurate information about the new residual: restart
       * it, p = z. */
      double prev = true_res;
      REFRESH(0);
      if (true_res < ORC_RESET_RATIO * prev)
        for (int i = 0; i < N; i++) p32[i] = poly ? w[0][i] : r32[i];
      continue;
    }
    if (s.iteration >= p->max_iter) {
      REFRESH(0);
      break;
    }
    int it = s.iteration;
    for (int i = 0; i < N; i++) q32[i] = op32_row(&o, p32, i);
    for (int i = 0; i < N; i++) f0[i] = (double)p32[i];
    for (int i = 0; i < N; i++) f1[i] = (double)q32[i];
    double pq = orc_rowblock_dot_tree_walk(o.nrb, o.starts, f0, f1, o.gmax, o.xcd);
    double alpha = s.rho / pq;
    if (h_pq) h_pq[it] = pq;
    if (h_alpha) h_alpha[it] = alpha;
    float a = (float)alpha, na = (float)(-alpha);
    for (int i = 0; i < N; i++) {
      y32[i] = y32[i] + a * p32[i];
      float r = r32[i] + na * q32[i];
      r32[i] = r;
      if (poly) w[deg % 3][i] = ck * r;
      f0[i] = (double)r * (double)r;
    }
    double rr = quad_tree(N, f0), rz = rr;
    if (poly) {
      poly_clenshaw(&o, r32, pl, w);
      for (int i = 0; i < N; i++) {
        f0[i] = (double)r32[i];
        f1[i] = (double)w[0][i];
      }
      rz = orc_rowblock_dot_tree_walk(o.nrb, o.starts, f0, f1, o.gmax, o.xcd);
    }
    double resid = sqrt(rr) / inner0;
    double beta = rz / s.rho;
    mp_scalars n2 = s;
    n2.iteration = it + 1;
    n2.rho = rz;
    n2.r2 = rr;
    n2.residual = resid;
    if (resid <= p->tol || rr <= delta2 * s.r2_ref) n2.done = 2;
    if (h_r2) h_r2[it + 1] = rr;
    if (h_rho) h_rho[it + 1] = rz;
    if (h_beta) h_beta[it] = beta;
    float b = (float)beta;
    for (int i = 0; i < N; i++) {
      float z = poly ? w[0][i] : r32[i];
      p32[i] = z + b * p32[i];
    }
    s = n2;
  }
#undef REFRESH
  res->iterations = s.iteration;
  res->updates = updates;
  res->true_residual = true_res;
  res->inner_residual = s.residual;
  free(y32), free(r32), free(p32), free(q32);
  free(q64), free(r64), free(f0), free(f1);
  for (int j = 0; j < 3; j++) free(w[j]);
  op32_free(&o);
  return 0;
}
