// C ABI of the MI355X CG backend (declared in include/cghip.h).
//
// The context owns one stream and every device buffer: the six CG vectors
// (CG::Vector, CG.h:22-35), the CRS and/or ELL matrix, the Jacobi diagonal,
// the reduction scratch and the device-resident solver state.  Each entry
// point validates its arguments on the host before anything is launched
// (indices in range, sizes consistent), so malformed input becomes an error
// code instead of an out-of-bounds access on the GPU.

#include "../../include/cghip.h"
#include "cghip_kernels.h"
#include "matrix_stats.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

using namespace cghip;

struct cghip_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;

  int n = 0;
  double *vec[6] = {};

  // CRS (MatrixDataCRS) + row blocks of the CSR-stream kernel
  bool has_crs = false;
  int crs_nnz = 0, crs_nrb = 0, crs_short = 0;
  int *crs_ptr = nullptr, *crs_index = nullptr, *crs_rb = nullptr;
  double *crs_value = nullptr;

  // ELL (MatrixDataELL), column-major
  bool has_ell = false;
  int ell_width = 0;
  int *ell_length = nullptr, *ell_index = nullptr;
  double *ell_data = nullptr;

  bool has_jac = false;
  double *jac = nullptr;

  int format = 0;
  int dot_mode = CGHIP_DOT_TREE;

  // reduction scratch
  // 4 x kPartialSlot: <p,q> | <r,r> | <r,z> | ||k - A x||^2
  double *partial = nullptr;
  // the FP64 fused loop's <p,q>: one partial per row block of the SpMV,
  // then their group sums (fold_groups) after them; sized at upload for
  // the larger of the uploaded matrices
  double *pq1 = nullptr;
  size_t pq1_cap = 0;
  double *dot_out = nullptr;

  // solver
  SolveState *d_state = nullptr;
  SolveState *h_state = nullptr;  // pinned
  double *h_scalar = nullptr;     // pinned
  double *d_hist = nullptr;
  int hist_cap = 0;
  int last_iterations = 0;

  // graph of `graph_batch` fused iterations
  hipGraphExec_t graph = nullptr;
  int graph_batch = 0, graph_format = 0, graph_precond = -1;

  // mixed precision: the FP32 copy of the CRS values; the scaled operator
  // D^-1/2 A D^-1/2 with implicit unit diagonal (off-diagonal CRS + its row
  // blocks, s = 1/sqrt(diag) in FP64); FP32 work vectors y32, r32, p32, q32;
  // the inner loop's state and graph
  bool has_fp32 = false;
  float *value32 = nullptr;
  bool has_scaled = false;
  int nnz_nd = 0, nrb_nd = 0, short_nd = 0;
  int *ptr_nd = nullptr, *index_nd = nullptr, *rb_nd = nullptr;
  float *value_nd = nullptr;
  double *scale64 = nullptr;
  // the scaled operator in column-major packed ELL when its padding is
  // small (and DIA does not apply)
  bool ell_nd = false;
  int ell_w_nd = 0;
  int *ell_len_nd = nullptr;
  uint64_t *ell_slot_nd = nullptr;  // packed (index | value bits << 32)
  // ... or in DIA when its off-diagonal entries lie on few diagonals
  // (found at upload: dia_ndiag = -1 if a row's columns are not ascending or
  // there are more than kDiaMax distinct offsets)
  int dia_ndiag = -1;
  int dia_off[kDiaMax] = {};
  bool dia_nd = false;
  float *dia_val = nullptr;
  uint16_t *dia_val16 = nullptr;  // FP16 copy for the polynomial's Clenshaw steps
  uint16_t *dia_up16 = nullptr;   // its upper diagonals when symmetric (Dia::up16), or null
  float *v32[4] = {};
  MixedState *d_mstate = nullptr;
  MixedState *h_mstate = nullptr;  // pinned
  hipGraphExec_t graph_mp = nullptr;
  int graph_mp_batch = 0, graph_mp_precond = -1, graph_mp_degree = 0, graph_mp_half = -1;

  // polynomial preconditioner: coefficients on the device, the three
  // Clenshaw work vectors (allocated by cghip_set_poly, freed with the FP32
  // data)
  bool has_poly = false;
  int poly_degree = 0;
  // the Clenshaw steps read the FP16 copy of the DIA values (default) or
  // the FP32 values (cghip_set_poly_values)
  bool poly_fp16 = true;
  PolyParams *d_poly = nullptr;
  float *w32[3] = {};

  std::vector<hipEvent_t> events;
};

namespace {

std::string g_create_err = "no error";
std::mutex g_create_mu;

int fail(cghip_ctx *c, int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  if (c)
    c->err = buf;
  else
    g_create_err = buf;
  return code;
}

#define HIPCHK(ctx, call)                                                                \
  do {                                                                                   \
    hipError_t e_ = (call);                                                              \
    if (e_ != hipSuccess)                                                                \
      return fail((ctx), CGHIP_ERR_HIP, "%s failed: %s", #call, hipGetErrorString(e_)); \
  } while (0)

template <typename T>
int dalloc(cghip_ctx *c, T **p, size_t count) {
  *p = nullptr;
  if (count == 0) count = 1;
  hipError_t e = hipMalloc((void **)p, count * sizeof(T));
  if (e != hipSuccess) {
    *p = nullptr;
    return fail(c, CGHIP_ERR_MEMORY, "hipMalloc of %zu bytes failed: %s", count * sizeof(T),
                hipGetErrorString(e));
  }
  return CGHIP_OK;
}

template <typename T>
void dfree(T *&p) {
  if (p) (void)hipFree((void *)p);
  p = nullptr;
}

void drop_graph(cghip_ctx *c) {
  if (c->graph) (void)hipGraphExecDestroy(c->graph);
  c->graph = nullptr;
  c->graph_batch = 0;
  c->graph_precond = -1;
  if (c->graph_mp) (void)hipGraphExecDestroy(c->graph_mp);
  c->graph_mp = nullptr;
  c->graph_mp_batch = 0;
  c->graph_mp_precond = -1;
  c->graph_mp_degree = 0;
}

void free_fp32(cghip_ctx *c) {
  dfree(c->value32);
  dfree(c->ptr_nd);
  dfree(c->index_nd);
  dfree(c->rb_nd);
  dfree(c->value_nd);
  dfree(c->scale64);
  dfree(c->ell_len_nd);
  dfree(c->ell_slot_nd);
  dfree(c->dia_val);
  dfree(c->dia_val16);
  dfree(c->dia_up16);
  c->dia_nd = false;
  c->ell_nd = false;
  c->ell_w_nd = 0;
  for (auto &v : c->v32) dfree(v);
  for (auto &v : c->w32) dfree(v);
  c->has_scaled = false;
  c->nnz_nd = c->nrb_nd = c->short_nd = 0;
  c->has_fp32 = false;
  c->has_poly = false;
  c->poly_degree = 0;
}

void free_crs(cghip_ctx *c) {
  dfree(c->crs_ptr);
  dfree(c->crs_index);
  dfree(c->crs_value);
  dfree(c->crs_rb);
  c->crs_short = 0;
  c->has_crs = false;
}

void free_ell(cghip_ctx *c) {
  dfree(c->ell_length);
  dfree(c->ell_index);
  dfree(c->ell_data);
  c->has_ell = false;
}

bool valid_vec(int id) { return id >= CGHIP_VEC_K && id <= CGHIP_VEC_Z; }

// Partial-sum slots: the SpMV's <p,q>; the vector kernels' two sums; the
// true residual of a reliable update.
Partials pq_of(cghip_ctx *c) { return Partials{c->partial, nullptr}; }
Partials pr_of(cghip_ctx *c) {
  return Partials{c->partial + kPartialSlot, c->partial + 2 * (size_t)kPartialSlot};
}
Partials pt_of(cghip_ctx *c) { return Partials{c->partial + 3 * (size_t)kPartialSlot, nullptr}; }

// Blocks of the FP64 SpMV + <p,q> kernel: one per row block (CRS) or
// 256-row chunk (ELL).
int spmv_blocks(cghip_ctx *c) {
  const int g = c->format == CGHIP_FORMAT_CRS ? c->crs_nrb : ell_chunks(c->n);
  return g > 0 ? g : 1;
}

// Room in pq1 for G partials and their fold_groups(G) group sums.
int ensure_pq1(cghip_ctx *c, int G) {
  const size_t need = (size_t)G + (size_t)fold_groups(G);
  if (c->pq1_cap >= need) return CGHIP_OK;
  dfree(c->pq1);
  c->pq1_cap = 0;
  if (int rc = dalloc(c, &c->pq1, need)) return rc;
  c->pq1_cap = need;
  return CGHIP_OK;
}

Crs<double> crs_of(cghip_ctx *c) {
  return Crs<double>{c->crs_ptr, c->crs_index, c->crs_value, c->crs_rb, c->n, c->crs_nrb, 0,
                     c->crs_short};
}

Ell<double> ell_of(cghip_ctx *c) {
  return Ell<double>{c->ell_length, c->ell_index, c->ell_data, c->n, c->ell_width};
}

Dia dia_of(cghip_ctx *c) {
  Dia d{c->dia_val, c->dia_val16, c->n, c->dia_ndiag > 0 ? c->dia_ndiag : 0, {}, 1, 0, 0,
        c->dia_up16};
  for (int k = 0; k < kDiaMax; k++) d.off[k] = c->dia_off[k];
  dia_strips(d);
  return d;
}

// The FP32 operator of the mixed inner loop: the scaled operator (unit
// diagonal implicit) in its DIA, ELL or CRS layout, or the unscaled FP32
// copy of A in CRS.
Op32 op32_of(cghip_ctx *c, bool scaled) {
  Op32 A{};
  const int n = c->n;
  if (scaled) {
    A.crs = Crs<float>{c->ptr_nd, c->index_nd, c->value_nd, c->rb_nd, n, c->nrb_nd, 1, c->short_nd};
    A.ell = EllPacked{c->ell_len_nd, c->ell_slot_nd, n, c->ell_w_nd, 1};
    A.dia = dia_of(c);
    A.kind = c->dia_nd ? kOpDia : c->ell_nd ? kOpEll : kOpCrs;
  } else {
    A.crs = Crs<float>{c->crs_ptr, c->crs_index, c->value32, c->crs_rb, n, c->crs_nrb, 0,
                       c->crs_short};
    A.kind = kOpCrs;
  }
  return A;
}

int check_ready(cghip_ctx *c) {
  if (!c) return CGHIP_ERR_ARG;
  if (c->n <= 0) return fail(c, CGHIP_ERR_STATE, "no vectors: call cghip_set_size first");
  if (c->format == CGHIP_FORMAT_CRS && !c->has_crs)
    return fail(c, CGHIP_ERR_STATE, "CRS format selected but no CRS matrix uploaded");
  if (c->format == CGHIP_FORMAT_ELL && !c->has_ell)
    return fail(c, CGHIP_ERR_STATE, "ELL format selected but no ELL matrix uploaded");
  if (c->format != CGHIP_FORMAT_CRS && c->format != CGHIP_FORMAT_ELL)
    return fail(c, CGHIP_ERR_STATE, "no matrix format selected");
  return CGHIP_OK;
}

hipError_t launch_matvec(cghip_ctx *c, const double *x, double *y) {
  if (c->format == CGHIP_FORMAT_CRS) return matvec_crs(crs_of(c), x, y, c->stream);
  return matvec_ell(ell_of(c), x, y, c->stream);
}

int sync(cghip_ctx *c) {
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return CGHIP_OK;
}

}  // namespace

void cghip::set_global_error(const char *msg) {
  std::lock_guard<std::mutex> lock(g_create_mu);
  g_create_err = msg;
}

extern "C" {

int cghip_geometry(int *block, int *vec_gmax, int *rows_max, int *nnz_cap, int *spmv_gmax,
                   int *ell_gmax, int *dia_max) {
  if (block) *block = kBlock;
  if (vec_gmax) *vec_gmax = kVecGMax;
  if (rows_max) *rows_max = kRowsMax;
  if (nnz_cap) *nnz_cap = kNnzCap;
  if (spmv_gmax) *spmv_gmax = kSpmvGMax;
  if (ell_gmax) *ell_gmax = kEllGMax;
  if (dia_max) *dia_max = kDiaMax;
  return CGHIP_OK;
}

const char *cghip_last_error(const cghip_ctx *ctx) {
  return ctx ? ctx->err.c_str() : g_create_err.c_str();
}

int cghip_create(int device, cghip_ctx **out) {
  if (!out) return fail(nullptr, CGHIP_ERR_ARG, "out is NULL");
  *out = nullptr;
  int count = 0;
  hipError_t e = hipGetDeviceCount(&count);
  if (e != hipSuccess || count <= 0)
    return fail(nullptr, CGHIP_ERR_HIP, "no HIP device available (%s)", hipGetErrorString(e));
  if (device < 0 || device >= count)
    return fail(nullptr, CGHIP_ERR_ARG, "device %d out of range [0, %d)", device, count);
  e = hipSetDevice(device);
  if (e != hipSuccess) return fail(nullptr, CGHIP_ERR_HIP, "hipSetDevice: %s", hipGetErrorString(e));
  cghip_ctx *c = new cghip_ctx;
  c->device = device;
  int rc = CGHIP_OK;
  if ((e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking)) != hipSuccess)
    rc = fail(nullptr, CGHIP_ERR_HIP, "hipStreamCreate: %s", hipGetErrorString(e));
  if (!rc && (e = noalloc_prepare(device)) != hipSuccess)
    rc = fail(nullptr, CGHIP_ERR_HIP, "scratch lookup: %s", hipGetErrorString(e));
  if (!rc) rc = dalloc(c, &c->partial, 4 * (size_t)kPartialSlot);
  if (!rc) rc = dalloc(c, &c->d_poly, 1);
  if (!rc) rc = dalloc(c, &c->dot_out, 8);
  if (!rc) rc = dalloc(c, &c->d_state, 1);
  if (!rc && (e = hipHostMalloc((void **)&c->h_state, sizeof(SolveState), 0)) != hipSuccess)
    rc = fail(nullptr, CGHIP_ERR_HIP, "hipHostMalloc: %s", hipGetErrorString(e));
  if (!rc && (e = hipHostMalloc((void **)&c->h_scalar, sizeof(double) * 8, 0)) != hipSuccess)
    rc = fail(nullptr, CGHIP_ERR_HIP, "hipHostMalloc: %s", hipGetErrorString(e));
  if (!rc && (e = hipDeviceSynchronize()) != hipSuccess)
    rc = fail(nullptr, CGHIP_ERR_HIP, "hipDeviceSynchronize: %s", hipGetErrorString(e));
  if (rc) {
    if (!c->err.empty()) g_create_err = c->err;
    cghip_destroy(c);
    return rc;
  }
  *out = c;
  return CGHIP_OK;
}

int cghip_destroy(cghip_ctx *c) {
  if (!c) return CGHIP_OK;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  drop_graph(c);
  for (auto &ev : c->events) (void)hipEventDestroy(ev);
  for (auto &v : c->vec) dfree(v);
  free_crs(c);
  free_fp32(c);
  free_ell(c);
  dfree(c->jac);
  dfree(c->partial);
  dfree(c->pq1);
  dfree(c->d_poly);
  dfree(c->dot_out);
  dfree(c->d_state);
  dfree(c->d_mstate);
  dfree(c->d_hist);
  if (c->h_state) (void)hipHostFree(c->h_state);
  if (c->h_mstate) (void)hipHostFree(c->h_mstate);
  if (c->h_scalar) (void)hipHostFree(c->h_scalar);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return CGHIP_OK;
}

int cghip_set_size(cghip_ctx *c, int n) {
  if (!c) return CGHIP_ERR_ARG;
  if (n <= 0) return fail(c, CGHIP_ERR_ARG, "n must be positive (got %d)", n);
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  drop_graph(c);
  if (c->n != n) {
    for (auto &v : c->vec) dfree(v);
    free_crs(c);
    free_fp32(c);
    free_ell(c);
    dfree(c->jac);
    c->has_jac = false;
    c->format = 0;
    c->n = 0;
    for (auto &v : c->vec)
      if (int rc = dalloc(c, &v, (size_t)n)) return rc;
    c->n = n;
  }
  for (auto &v : c->vec) HIPCHK(c, hipMemsetAsync(v, 0, sizeof(double) * (size_t)n, c->stream));
  return sync(c);
}

int cghip_upload_crs(cghip_ctx *c, int n, int nnz, const int32_t *ptr, const int32_t *index,
                     const double *value) {
  if (!c) return CGHIP_ERR_ARG;
  if (n <= 0 || nnz < 0 || !ptr || (nnz > 0 && (!index || !value)))
    return fail(c, CGHIP_ERR_ARG, "invalid CRS arguments (n=%d, nnz=%d)", n, nnz);
  if (ptr[0] != 0 || ptr[n] != nnz)
    return fail(c, CGHIP_ERR_ARG, "CRS ptr must start at 0 and end at nnz");
  for (int i = 0; i < n; i++)
    if (ptr[i + 1] < ptr[i]) return fail(c, CGHIP_ERR_ARG, "CRS ptr decreases at row %d", i);
  for (int e = 0; e < nnz; e++)
    if (index[e] < 0 || index[e] >= n)
      return fail(c, CGHIP_ERR_ARG, "CRS column index %d out of range at entry %d", index[e], e);
  // distinct off-diagonal offsets j - i (ascending), for the DIA layout
  int ndiag = 0;
  int off[kDiaMax + 1];
  for (int i = 0; i < n && ndiag >= 0; i++)
    for (int e = ptr[i]; e < ptr[i + 1]; e++) {
      if (e > ptr[i] && index[e] <= index[e - 1]) {
        ndiag = -1;  // columns not ascending: the DIA order would differ
        break;
      }
      const int o = index[e] - i;
      if (o == 0) continue;
      int d = 0;
      while (d < ndiag && off[d] < o) d++;
      if (d < ndiag && off[d] == o) continue;
      if (ndiag == kDiaMax) {
        ndiag = -1;
        break;
      }
      for (int k = ndiag; k > d; k--) off[k] = off[k - 1];
      off[d] = o;
      ndiag++;
    }
  if (c->n != n)
    if (int rc = cghip_set_size(c, n)) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  drop_graph(c);
  free_crs(c);
  free_fp32(c);
  std::vector<int> rb((size_t)n + 1);
  const int nrb = crs_rowblocks(n, ptr, rb.data());
  int rc = dalloc(c, &c->crs_ptr, (size_t)n + 1);
  if (!rc) rc = dalloc(c, &c->crs_index, (size_t)nnz);
  if (!rc) rc = dalloc(c, &c->crs_value, (size_t)nnz);
  if (!rc) rc = dalloc(c, &c->crs_rb, (size_t)nrb + 1);
  if (rc) {
    free_crs(c);
    return rc;
  }
  HIPCHK(c, hipMemcpy(c->crs_ptr, ptr, sizeof(int) * ((size_t)n + 1), hipMemcpyHostToDevice));
  if (nnz > 0) {
    HIPCHK(c, hipMemcpy(c->crs_index, index, sizeof(int) * (size_t)nnz, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->crs_value, value, sizeof(double) * (size_t)nnz, hipMemcpyHostToDevice));
  }
  HIPCHK(c, hipMemcpy(c->crs_rb, rb.data(), sizeof(int) * ((size_t)nrb + 1), hipMemcpyHostToDevice));
  if (int e = ensure_pq1(c, nrb > 0 ? nrb : 1)) {
    free_crs(c);
    return e;
  }
  c->crs_nnz = nnz;
  c->crs_nrb = nrb;
  c->crs_short = crs_short_rows(n, ptr, nrb);
  c->dia_ndiag = ndiag;
  for (int d = 0; d < kDiaMax; d++) c->dia_off[d] = d < ndiag ? off[d] : 0;
  c->has_crs = true;
  if (!c->format) c->format = CGHIP_FORMAT_CRS;
  return CGHIP_OK;
}

int cghip_upload_ell(cghip_ctx *c, int n, int width, const int32_t *length, const int32_t *index,
                     const double *data) {
  if (!c) return CGHIP_ERR_ARG;
  if (n <= 0 || width < 0 || !length || (width > 0 && (!index || !data)))
    return fail(c, CGHIP_ERR_ARG, "invalid ELL arguments (n=%d, width=%d)", n, width);
  const int64_t slots = (int64_t)n * width;
  for (int i = 0; i < n; i++) {
    if (length[i] < 0 || length[i] > width)
      return fail(c, CGHIP_ERR_ARG, "ELL length[%d] = %d outside [0, %d]", i, length[i], width);
    for (int j = 0; j < length[i]; j++) {
      const int col = index[(int64_t)j * n + i];
      if (col < 0 || col >= n)
        return fail(c, CGHIP_ERR_ARG, "ELL column index %d out of range (row %d, slot %d)", col, i, j);
    }
  }
  if (c->n != n)
    if (int rc = cghip_set_size(c, n)) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  drop_graph(c);
  free_ell(c);
  int rc = dalloc(c, &c->ell_length, (size_t)n);
  if (!rc) rc = dalloc(c, &c->ell_index, (size_t)slots);
  if (!rc) rc = dalloc(c, &c->ell_data, (size_t)slots);
  if (rc) {
    free_ell(c);
    return rc;
  }
  HIPCHK(c, hipMemcpy(c->ell_length, length, sizeof(int) * (size_t)n, hipMemcpyHostToDevice));
  if (slots > 0) {
    HIPCHK(c, hipMemcpy(c->ell_index, index, sizeof(int) * (size_t)slots, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->ell_data, data, sizeof(double) * (size_t)slots, hipMemcpyHostToDevice));
  }
  if (int e = ensure_pq1(c, std::max(ell_chunks(n), 1))) {
    free_ell(c);
    return e;
  }
  c->ell_width = width;
  c->has_ell = true;
  if (!c->format) c->format = CGHIP_FORMAT_ELL;
  return CGHIP_OK;
}

int cghip_upload_jacobi(cghip_ctx *c, const double *C) {
  if (!c || !C) return c ? fail(c, CGHIP_ERR_ARG, "C is NULL") : CGHIP_ERR_ARG;
  if (c->n <= 0) return fail(c, CGHIP_ERR_STATE, "no size: upload a matrix first");
  HIPCHK(c, hipStreamSynchronize(c->stream));
  drop_graph(c);
  if (!c->jac)
    if (int rc = dalloc(c, &c->jac, (size_t)c->n)) return rc;
  HIPCHK(c, hipMemcpy(c->jac, C, sizeof(double) * (size_t)c->n, hipMemcpyHostToDevice));
  free_fp32(c);  // the FP32 diagonal is stale
  c->has_jac = true;
  return CGHIP_OK;
}

int cghip_set_format(cghip_ctx *c, int format) {
  if (!c) return CGHIP_ERR_ARG;
  if (format == CGHIP_FORMAT_CRS && !c->has_crs)
    return fail(c, CGHIP_ERR_STATE, "no CRS matrix uploaded");
  if (format == CGHIP_FORMAT_ELL && !c->has_ell)
    return fail(c, CGHIP_ERR_STATE, "no ELL matrix uploaded");
  if (format != CGHIP_FORMAT_CRS && format != CGHIP_FORMAT_ELL)
    return fail(c, CGHIP_ERR_ARG, "unsupported matrix format %d", format);
  if (c->format != format) drop_graph(c);
  c->format = format;
  return CGHIP_OK;
}

int cghip_set_dot_mode(cghip_ctx *c, int mode) {
  if (!c) return CGHIP_ERR_ARG;
  if (mode != CGHIP_DOT_TREE && mode != CGHIP_DOT_SERIAL)
    return fail(c, CGHIP_ERR_ARG, "unknown dot mode %d", mode);
  c->dot_mode = mode;
  return CGHIP_OK;
}

int cghip_upload_vector(cghip_ctx *c, int id, const double *host) {
  if (!c) return CGHIP_ERR_ARG;
  if (!valid_vec(id) || !host) return fail(c, CGHIP_ERR_ARG, "invalid vector id %d or NULL", id);
  if (c->n <= 0) return fail(c, CGHIP_ERR_STATE, "no size set");
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipMemcpyAsync(c->vec[id], host, sizeof(double) * (size_t)c->n, hipMemcpyHostToDevice,
                           c->stream));
  return sync(c);
}

int cghip_download_vector(cghip_ctx *c, int id, double *host) {
  if (!c) return CGHIP_ERR_ARG;
  if (!valid_vec(id) || !host) return fail(c, CGHIP_ERR_ARG, "invalid vector id %d or NULL", id);
  if (c->n <= 0) return fail(c, CGHIP_ERR_STATE, "no size set");
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipMemcpyAsync(host, c->vec[id], sizeof(double) * (size_t)c->n, hipMemcpyDeviceToHost,
                           c->stream));
  return sync(c);
}

int cghip_vector_pointer(cghip_ctx *c, int id, double **out) {
  if (!c) return CGHIP_ERR_ARG;
  if (!valid_vec(id) || !out) return fail(c, CGHIP_ERR_ARG, "invalid vector id %d or NULL", id);
  if (c->n <= 0) return fail(c, CGHIP_ERR_STATE, "no size set");
  *out = c->vec[id];
  return CGHIP_OK;
}

// ---- hooks ---------------------------------------------------------------

int cghip_cpy(cghip_ctx *c, int dst, int src) {
  if (!c) return CGHIP_ERR_ARG;
  if (!valid_vec(dst) || !valid_vec(src)) return fail(c, CGHIP_ERR_ARG, "invalid vector id");
  if (c->n <= 0) return fail(c, CGHIP_ERR_STATE, "no size set");
  HIPCHK(c, hipMemcpyAsync(c->vec[dst], c->vec[src], sizeof(double) * (size_t)c->n,
                           hipMemcpyDeviceToDevice, c->stream));
  return sync(c);
}

int cghip_matvec(cghip_ctx *c, int x, int y) {
  if (int rc = check_ready(c)) return rc;
  if (!valid_vec(x) || !valid_vec(y) || x == y)
    return fail(c, CGHIP_ERR_ARG, "invalid vectors for matvec (%d -> %d)", x, y);
  HIPCHK(c, launch_matvec(c, c->vec[x], c->vec[y]));
  return sync(c);
}

int cghip_axpy(cghip_ctx *c, double a, int x, int y) {
  if (!c) return CGHIP_ERR_ARG;
  if (!valid_vec(x) || !valid_vec(y)) return fail(c, CGHIP_ERR_ARG, "invalid vector id");
  if (c->n <= 0) return fail(c, CGHIP_ERR_STATE, "no size set");
  HIPCHK(c, axpy(c->n, a, c->vec[x], c->vec[y], c->stream));
  return sync(c);
}

int cghip_xpay(cghip_ctx *c, int x, double a, int y) {
  if (!c) return CGHIP_ERR_ARG;
  if (!valid_vec(x) || !valid_vec(y)) return fail(c, CGHIP_ERR_ARG, "invalid vector id");
  if (c->n <= 0) return fail(c, CGHIP_ERR_STATE, "no size set");
  HIPCHK(c, xpay(c->n, c->vec[x], a, c->vec[y], c->stream));
  return sync(c);
}

int cghip_vector_dot(cghip_ctx *c, int a, int b, double *out) {
  if (!c) return CGHIP_ERR_ARG;
  if (!valid_vec(a) || !valid_vec(b) || !out) return fail(c, CGHIP_ERR_ARG, "invalid arguments");
  if (c->n <= 0) return fail(c, CGHIP_ERR_STATE, "no size set");
  if (c->dot_mode == CGHIP_DOT_SERIAL)
    HIPCHK(c, dot_serial(c->n, c->vec[a], c->vec[b], c->dot_out, c->stream));
  else
    HIPCHK(c, dot_tree(c->n, c->vec[a], c->vec[b], pr_of(c), c->dot_out, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->h_scalar, c->dot_out, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  if (int rc = sync(c)) return rc;
  *out = c->h_scalar[0];
  return CGHIP_OK;
}

int cghip_apply_preconditioner(cghip_ctx *c, int x, int y) {
  if (!c) return CGHIP_ERR_ARG;
  if (!valid_vec(x) || !valid_vec(y)) return fail(c, CGHIP_ERR_ARG, "invalid vector id");
  if (!c->has_jac) return fail(c, CGHIP_ERR_STATE, "no Jacobi preconditioner uploaded");
  HIPCHK(c, jacobi(c->n, c->jac, c->vec[x], c->vec[y], c->stream));
  return sync(c);
}

// ---- fused solver --------------------------------------------------------

namespace {

// One fused iteration.  With `ev`, events ev[0..3] bracket the three kernels.
int enqueue_iteration(cghip_ctx *c, const double *C, History h, hipEvent_t *ev) {
  double *p = c->vec[CGHIP_VEC_P], *q = c->vec[CGHIP_VEC_Q], *r = c->vec[CGHIP_VEC_R],
         *x = c->vec[CGHIP_VEC_X];
  if (ev) HIPCHK(c, hipEventRecord(ev[0], c->stream));
  const int G = spmv_blocks(c);
  // G <= 256: the update folds the SpMV's partials itself (no group fold)
  const bool direct = G <= kBlock;
  const Partials pq1{c->pq1, nullptr}, pq2{c->pq1 + G, nullptr};
  double *grp = direct ? nullptr : pq2.a;
  if (c->format == CGHIP_FORMAT_CRS)
    HIPCHK(c, pcg_matvec_dot_crs(crs_of(c), p, q, pq1, grp, c->d_state, c->stream));
  else
    HIPCHK(c, pcg_matvec_dot_ell(ell_of(c), p, q, pq1, grp, c->d_state, c->stream));
  if (ev) HIPCHK(c, hipEventRecord(ev[1], c->stream));
  HIPCHK(c, pcg_update(c->n, direct ? -G : fold_groups(G), r, q, C, direct ? pq1 : pq2, pr_of(c),
                       c->d_state, h, c->stream));
  if (ev) HIPCHK(c, hipEventRecord(ev[2], c->stream));
  HIPCHK(c, pcg_xpay(c->n, x, p, r, C, pr_of(c), c->d_state, h, c->stream));
  if (ev) HIPCHK(c, hipEventRecord(ev[3], c->stream));
  return CGHIP_OK;
}

int ensure_events(cghip_ctx *c, size_t count) {
  while (c->events.size() < count) {
    hipEvent_t e;
    HIPCHK(c, hipEventCreate(&e));
    c->events.push_back(e);
  }
  return CGHIP_OK;
}

// Milliseconds between two completed events (0 if the runtime cannot tell;
// the error is cleared so that it is not reported by a later launch check).
float elapsed(hipEvent_t a, hipEvent_t b) {
  float ms = 0.f;
  if (hipEventElapsedTime(&ms, a, b) != hipSuccess) {
    (void)hipGetLastError();
    ms = 0.f;
  }
  return ms;
}

// History buffers for a solve of at most max_iter iterations: min(max_iter
// + 1, kHistCap) entries per array, in 64-bit arithmetic (max_iter may be
// INT_MAX), reallocated only when a larger one is needed; the kernels write
// entry i only while i < h->cap.
int ensure_history(cghip_ctx *c, int max_iter, History *h) {
  const int want = (int)std::min<int64_t>((int64_t)max_iter + 1, (int64_t)kHistCap);
  if (c->hist_cap < want) {
    drop_graph(c);
    dfree(c->d_hist);
    c->hist_cap = 0;
    if (int rc = dalloc(c, &c->d_hist, 5 * (size_t)want)) return rc;
    c->hist_cap = want;
  }
  const size_t hc = (size_t)c->hist_cap;
  *h = History{c->d_hist,          c->d_hist + hc,     c->d_hist + 2 * hc,
               c->d_hist + 3 * hc, c->d_hist + 4 * hc, c->hist_cap};
  return CGHIP_OK;
}

}  // namespace

int cghip_solve(cghip_ctx *c, const cghip_solve_opts *o, cghip_solve_info *info) {
  if (int rc = check_ready(c)) return rc;
  if (!o || !info) return fail(c, CGHIP_ERR_ARG, "opts/info is NULL");
  if (o->max_iter <= 0) return fail(c, CGHIP_ERR_ARG, "max_iter must be positive");
  if (!(o->tol > 0)) return fail(c, CGHIP_ERR_ARG, "tol must be positive");
  if (o->precond != CGHIP_PRECOND_NONE && o->precond != CGHIP_PRECOND_JACOBI)
    return fail(c, CGHIP_ERR_ARG, "unsupported preconditioner %d", o->precond);
  if (o->precond == CGHIP_PRECOND_JACOBI && !c->has_jac)
    return fail(c, CGHIP_ERR_STATE, "Jacobi requested but no diagonal uploaded");
  HIPCHK(c, hipSetDevice(c->device));
  std::memset(info, 0, sizeof(*info));

  History h;
  if (int rc = ensure_history(c, o->max_iter, &h)) return rc;
  const double *C = o->precond == CGHIP_PRECOND_JACOBI ? c->jac : nullptr;

  SolveState init{};
  init.tol = o->tol;
  init.max_iter = o->max_iter;
  *c->h_state = init;
  HIPCHK(c, hipMemcpyAsync(c->d_state, c->h_state, sizeof(SolveState), hipMemcpyHostToDevice,
                           c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));

  if (o->profile != CGHIP_PROFILE_NONE && o->profile != CGHIP_PROFILE_KERNELS)
    return fail(c, CGHIP_ERR_ARG, "unknown profile mode %d", o->profile);
  const int mode = o->profile;
  // Iterations per graph launch; never more than max_iter (the kernels
  // after CG_MAX_ITER exit at once but still cost a launch each, e.g. the
  // 10-step solves of the learned preconditioner's features).
  int batch = o->batch > 0 ? o->batch : (c->n >= (1 << 20) ? 8 : 32);
  if (o->batch <= 0) batch = std::min(batch, o->max_iter);
  if (mode == CGHIP_PROFILE_KERNELS) batch = 1;
  // events[0..1] bracket the set-up; events[2..5] the kernels of an iteration.
  if (mode != CGHIP_PROFILE_NONE && ensure_events(c, 6)) return CGHIP_ERR_HIP;
  hipEvent_t *ev = mode != CGHIP_PROFILE_NONE ? c->events.data() : nullptr;

  auto read_state = [&]() -> int {
    HIPCHK(c, hipMemcpyAsync(c->h_state, c->d_state, sizeof(SolveState), hipMemcpyDeviceToHost,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return CGHIP_OK;
  };
  // Charges the kernels of one iteration timed by ev[2..5]; its p update was
  // skipped if the solve converged there.
  auto charge = [&](bool converged) {
    hipEvent_t *e = ev + 2;
    info->ms_matvec += elapsed(e[0], e[1]);
    info->ms_update += elapsed(e[1], e[2]);
    info->n_matvec++;
    info->n_update++;
    if (!converged) {
      info->ms_xpay += elapsed(e[2], e[3]);
      info->n_xpay++;
    }
  };

  const auto t0 = std::chrono::steady_clock::now();
  // r = A x; r = k - r; p = C r; r2 = <r,r>; rho = <p,r>
  if (ev) HIPCHK(c, hipEventRecord(ev[0], c->stream));
  HIPCHK(c, launch_matvec(c, c->vec[CGHIP_VEC_X], c->vec[CGHIP_VEC_R]));
  HIPCHK(c, pcg_init(c->n, c->vec[CGHIP_VEC_K], c->vec[CGHIP_VEC_R], c->vec[CGHIP_VEC_P], C,
                     pr_of(c), c->d_state, h, c->stream));
  if (ev) HIPCHK(c, hipEventRecord(ev[1], c->stream));
  if (int rc = read_state()) return rc;
  if (ev) {
    info->ms_init = elapsed(ev[0], ev[1]);
    info->n_init = 1;
  }

  if (mode == CGHIP_PROFILE_KERNELS) {
    for (;;) {
      const Scalars &nx = c->h_state->next;
      if (nx.done || nx.iteration >= o->max_iter) break;
      if (int rc = enqueue_iteration(c, C, h, ev + 2)) return rc;
      if (int rc = read_state()) return rc;
      charge(c->h_state->next.done != 0);
    }
  } else {
    if (!c->graph || c->graph_batch != batch || c->graph_format != c->format ||
        c->graph_precond != o->precond) {
      drop_graph(c);
      hipGraph_t g = nullptr;
      HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
      int rc = CGHIP_OK;
      for (int i = 0; i < batch && !rc; i++) rc = enqueue_iteration(c, C, h, nullptr);
      hipError_t e = hipStreamEndCapture(c->stream, &g);
      if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
      }
      if (e != hipSuccess)
        return fail(c, CGHIP_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
      e = hipGraphInstantiate(&c->graph, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (e != hipSuccess) {
        c->graph = nullptr;
        return fail(c, CGHIP_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
      }
      c->graph_batch = batch;
      c->graph_format = c->format;
      c->graph_precond = o->precond;
    }
    // Every batch advances the device counter by `batch` iterations until
    // CG_MAX_ITER stops it, so this bound is never reached by a correct run.
    const long max_batches = (long)o->max_iter / batch + 3;
    for (long nb = 0;; nb++) {
      if (nb >= max_batches)
        return fail(c, CGHIP_ERR_STATE, "solver did not stop after %ld batches", nb);
      HIPCHK(c, hipGraphLaunch(c->graph, c->stream));
      if (int rc = read_state()) return rc;
      if (c->h_state->cur.done || c->h_state->next.done) break;
    }
  }
  const auto t1 = std::chrono::steady_clock::now();
  const Scalars &s = c->h_state->next;
  c->last_iterations = s.iteration;
  info->iterations = s.iteration;
  info->residual = s.residual;
  info->r2 = s.r2;
  info->wall_s = std::chrono::duration<double>(t1 - t0).count();
  return CGHIP_OK;
}

// ---- mixed precision -----------------------------------------------------

int cghip_prepare_fp32(cghip_ctx *c) {
  if (!c) return CGHIP_ERR_ARG;
  if (!c->has_crs) return fail(c, CGHIP_ERR_STATE, "mixed precision needs a CRS matrix");
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  drop_graph(c);
  free_fp32(c);
  const int n = c->n;
  int rc = dalloc(c, &c->value32, (size_t)c->crs_nnz);
  for (auto &v : c->v32)
    if (!rc) rc = dalloc(c, &v, (size_t)n);
  if (!rc) rc = dalloc(c, &c->scale64, (size_t)n);
  if (!rc && !c->d_mstate) rc = dalloc(c, &c->d_mstate, 1);
  if (!rc && !c->h_mstate) {
    hipError_t e = hipHostMalloc((void **)&c->h_mstate, sizeof(MixedState), 0);
    if (e != hipSuccess) rc = fail(c, CGHIP_ERR_HIP, "hipHostMalloc: %s", hipGetErrorString(e));
  }
  if (rc) {
    free_fp32(c);
    return rc;
  }
  HIPCHK(c, convert_f64_f32(c->crs_nnz, c->crs_value, c->value32, c->stream));
  for (auto &v : c->v32) HIPCHK(c, hipMemsetAsync(v, 0, sizeof(float) * (size_t)n, c->stream));
  // Scaled operator: needs exactly one positive diagonal entry per row.
  int *d_bad = reinterpret_cast<int *>(c->dot_out);
  HIPCHK(c, hipMemsetAsync(d_bad, 0, sizeof(int), c->stream));
  HIPCHK(c, scaled_diag(crs_of(c), c->scale64, d_bad, c->stream));
  int bad = 0;
  HIPCHK(c, hipMemcpyAsync(&bad, d_bad, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  if (int r = sync(c)) return r;
  if (!bad) {
    // row i keeps ptr[i+1] - ptr[i] - 1 entries: the off-diagonal row
    // pointer is ptr[i] - i
    std::vector<int> ptr(n + 1);
    HIPCHK(c, hipMemcpy(ptr.data(), c->crs_ptr, sizeof(int) * (n + 1), hipMemcpyDeviceToHost));
    for (int i = 0; i <= n; i++) ptr[i] -= i;
    c->nnz_nd = ptr[n];
    std::vector<int> starts(n + 1);
    c->nrb_nd = crs_rowblocks(n, ptr.data(), starts.data());
    c->short_nd = crs_short_rows(n, ptr.data(), c->nrb_nd);
    rc = dalloc(c, &c->ptr_nd, (size_t)n + 1);
    if (!rc) rc = dalloc(c, &c->index_nd, (size_t)c->nnz_nd);
    if (!rc) rc = dalloc(c, &c->value_nd, (size_t)c->nnz_nd);
    if (!rc) rc = dalloc(c, &c->rb_nd, (size_t)c->nrb_nd + 1);
    if (rc) {
      free_fp32(c);
      return rc;
    }
    HIPCHK(c, hipMemcpy(c->ptr_nd, ptr.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->rb_nd, starts.data(), sizeof(int) * (c->nrb_nd + 1),
                        hipMemcpyHostToDevice));
    HIPCHK(c, scaled_copy(crs_of(c), c->scale64, c->index_nd, c->value_nd, c->stream));
    int w = 0;
    for (int i = 0; i < n; i++) w = std::max(w, ptr[i + 1] - ptr[i]);
    if (c->dia_ndiag > 0 && dia_pays(c->dia_ndiag, n, c->nnz_nd)) {
      const size_t slots = (size_t)c->dia_ndiag * (size_t)n;
      rc = dalloc(c, &c->dia_val, slots);
      if (!rc) rc = dalloc(c, &c->dia_val16, 2 * (size_t)half_words(c->dia_ndiag) * (size_t)n);
      if (rc) {
        free_fp32(c);
        return rc;
      }
      int *d_over = reinterpret_cast<int *>(c->dot_out);
      HIPCHK(c, hipMemsetAsync(d_over, 0, sizeof(int), c->stream));
      HIPCHK(c, scaled_dia(crs_of(c), c->scale64, dia_of(c), c->dia_val, c->dia_val16, d_over,
                           c->stream));
      int over = 0;
      HIPCHK(c, hipMemcpyAsync(&over, d_over, sizeof(int), hipMemcpyDeviceToHost, c->stream));
      if (int r = sync(c)) return r;
      c->dia_nd = true;
      if (over) {  // no FP16 copy possible: the next layout instead
        dfree(c->dia_val);
        dfree(c->dia_val16);
        c->dia_nd = false;
      } else if (c->dia_ndiag == 2 * kSymPairs && dia_mirrored(dia_of(c))) {
        // symmetric FP16 storage, kept only if every value equals its
        // transpose's bit for bit
        rc = dalloc(c, &c->dia_up16, (size_t)kSymPairs * ((size_t)n + kSymPad));
        if (rc) {
          free_fp32(c);
          return rc;
        }
        int *d_mis = reinterpret_cast<int *>(c->dot_out);
        HIPCHK(c, hipMemsetAsync(d_mis, 0, sizeof(int), c->stream));
        Dia d = dia_of(c);
        d.up16 = nullptr;
        HIPCHK(c, dia_upper16(d, c->dia_up16, d_mis, c->stream));
        int mis = 0;
        HIPCHK(c, hipMemcpyAsync(&mis, d_mis, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        if (int r = sync(c)) return r;
        if (mis) dfree(c->dia_up16);
      }
    }
    if (!c->dia_nd && ell_pays(w, n, c->nnz_nd)) {
      const size_t slots = (size_t)w * (size_t)n;
      rc = dalloc(c, &c->ell_len_nd, (size_t)n);
      if (!rc) rc = dalloc(c, &c->ell_slot_nd, slots);
      if (rc) {
        free_fp32(c);
        return rc;
      }
      HIPCHK(c, scaled_ell(crs_of(c), c->scale64, c->ell_len_nd, c->ell_slot_nd, c->stream));
      c->ell_nd = true;
      c->ell_w_nd = w;
    }
    if (int r = sync(c)) return r;
    c->has_scaled = true;
  }
  c->has_fp32 = true;
  return CGHIP_OK;
}

int cghip_set_poly_values(cghip_ctx *c, int bits) {
  if (!c) return CGHIP_ERR_ARG;
  if (bits != 16 && bits != 32)
    return fail(c, CGHIP_ERR_ARG, "value precision must be 16 or 32 bits (got %d)", bits);
  if (c->poly_fp16 != (bits == 16)) {
    c->poly_fp16 = bits == 16;
    c->has_poly = false;  // its positivity check depended on the values
  }
  return CGHIP_OK;
}

int cghip_fp32_layout(const cghip_ctx *c, int *layout) {
  if (!c || !layout) return CGHIP_ERR_ARG;
  *layout = !c->has_scaled ? -1
            : c->dia_nd    ? CGHIP_LAYOUT_DIA
            : c->ell_nd    ? CGHIP_LAYOUT_ELL
                           : CGHIP_LAYOUT_CRS;
  return CGHIP_OK;
}

int cghip_poly_values_layout(const cghip_ctx *c, int *layout) {
  if (!c || !layout) return CGHIP_ERR_ARG;
  *layout = !c->has_scaled                  ? -1
            : !c->dia_nd || !c->poly_fp16   ? CGHIP_POLY_VALUES_FP32
            : c->dia_up16                   ? CGHIP_POLY_VALUES_FP16_SYM
                                            : CGHIP_POLY_VALUES_FP16;
  return CGHIP_OK;
}

int cghip_set_poly(cghip_ctx *c, int degree, const double *coeffs, double b) {
  if (!c) return CGHIP_ERR_ARG;
  if (!c->has_fp32) return fail(c, CGHIP_ERR_STATE, "call cghip_prepare_fp32 first");
  if (!c->has_scaled)
    return fail(c, CGHIP_ERR_STATE,
                "the polynomial preconditioner needs one positive diagonal entry per row");
  if (degree < 1 || degree > kPolyMaxDegree)
    return fail(c, CGHIP_ERR_ARG, "degree %d outside [1, %d]", degree, kPolyMaxDegree);
  if (!coeffs) return fail(c, CGHIP_ERR_ARG, "coeffs is NULL");
  if (!(b > 0) || !std::isfinite(b) || !std::isfinite((float)(2.0 / b)))
    return fail(c, CGHIP_ERR_ARG, "spectrum bound b = %g must be positive and finite", b);
  PolyParams pp{};
  for (int j = 0; j <= degree; j++) {
    pp.c[j] = (float)coeffs[j];
    if (!std::isfinite(coeffs[j]) || !std::isfinite(pp.c[j]))
      return fail(c, CGHIP_ERR_ARG, "coefficient %d is not a finite float", j);
  }
  pp.two_over_b = (float)(2.0 / b);
  pp.degree = degree;
  // p > 0 on [0, b] (the float-rounded coefficients), evaluated in FP64 by
  // Clenshaw's recurrence at y = 2 l / b - 1 on a uniform grid.  With the
  // DIA layout the Clenshaw steps apply the FP16-rounded operator, whose
  // entries differ by at most 2^-11 relative (|a_ij| <= 1 after scaling), so
  // its spectrum lies in [-2^-10 b, (1 + 2^-10) b]: p is checked there.
  constexpr int kGrid = 4096;
  const double ext = c->dia_nd && c->poly_fp16 ? 2.0 * std::ldexp(1.0, -10) : 0.0;  // in y units
  for (int i = 0; i < kGrid; i++) {
    const double y = -1.0 - ext + 2.0 * (1.0 + ext) * i / (kGrid - 1);
    double b1 = 0.0, b2 = 0.0;
    for (int j = degree; j >= 1; j--) {
      const double t = (double)pp.c[j] + 2.0 * y * b1 - b2;
      b2 = b1;
      b1 = t;
    }
    const double pv = (double)pp.c[0] + y * b1 - b2;
    if (!(pv > 0))
      return fail(c, CGHIP_ERR_ARG, "p(%g) = %g is not positive: the preconditioner would not be SPD",
                  (y + 1.0) * 0.5 * b, pv);
  }
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (auto &v : c->w32)
    if (!v)
      if (int rc = dalloc(c, &v, (size_t)c->n)) return rc;
  HIPCHK(c, hipMemcpyAsync(c->d_poly, &pp, sizeof(PolyParams), hipMemcpyHostToDevice, c->stream));
  if (int r = sync(c)) return r;
  c->poly_degree = degree;
  c->has_poly = true;
  return CGHIP_OK;
}

int cghip_solve_mixed(cghip_ctx *c, const cghip_mixed_opts *o, cghip_mixed_info *info) {
  if (!c) return CGHIP_ERR_ARG;
  if (!o || !info) return fail(c, CGHIP_ERR_ARG, "opts/info is NULL");
  if (!c->has_crs || c->format != CGHIP_FORMAT_CRS)
    return fail(c, CGHIP_ERR_STATE, "mixed precision needs the CRS format selected");
  if (!c->has_fp32) return fail(c, CGHIP_ERR_STATE, "call cghip_prepare_fp32 first");
  if (o->max_iter <= 0) return fail(c, CGHIP_ERR_ARG, "max_iter must be positive");
  if (!(o->tol > 0)) return fail(c, CGHIP_ERR_ARG, "tol must be positive");
  if (!(o->delta >= 0 && o->delta < 1)) return fail(c, CGHIP_ERR_ARG, "delta must be in [0, 1)");
  if (o->precond != CGHIP_PRECOND_NONE && o->precond != CGHIP_PRECOND_JACOBI &&
      o->precond != CGHIP_PRECOND_POLY)
    return fail(c, CGHIP_ERR_ARG, "unsupported preconditioner %d", o->precond);
  if (o->precond == CGHIP_PRECOND_JACOBI && !c->has_scaled)
    return fail(c, CGHIP_ERR_STATE, "Jacobi needs one positive diagonal entry per row");
  if (o->precond == CGHIP_PRECOND_POLY && !c->has_poly)
    return fail(c, CGHIP_ERR_STATE, "polynomial requested: call cghip_set_poly first");
  if (o->profile != CGHIP_PROFILE_NONE && o->profile != CGHIP_PROFILE_KERNELS)
    return fail(c, CGHIP_ERR_ARG, "unknown profile mode %d", o->profile);
  HIPCHK(c, hipSetDevice(c->device));
  std::memset(info, 0, sizeof(*info));
  History h;
  if (int rc = ensure_history(c, o->max_iter, &h)) return rc;

  const int n = c->n;
  const bool poly = o->precond == CGHIP_PRECOND_POLY;
  const int degree = poly ? c->poly_degree : 0;
  // Jacobi and the polynomial run on the scaled operator (unit diagonal
  // implicit); no preconditioner runs on the FP32 copy of A.
  const bool scaled = o->precond != CGHIP_PRECOND_NONE;
  const double *scale = scaled ? c->scale64 : nullptr;
  float *const w[3] = {c->w32[0], c->w32[1], c->w32[2]};
  float *y32 = c->v32[0], *r32 = c->v32[1], *p32 = c->v32[2], *q32 = c->v32[3];
  double *x64 = c->vec[CGHIP_VEC_X], *k64 = c->vec[CGHIP_VEC_K], *r64 = c->vec[CGHIP_VEC_R],
         *q64 = c->vec[CGHIP_VEC_Q];
  const Op32 A32 = op32_of(c, scaled);
  const int gpq = op32_grid(A32);
  // slots: <p,q> (SpMV grid), <r,r> (quads grid), <r,z> (SpMV grid, the
  // polynomial only: otherwise z = r), ||k - A x||^2 (pairs grid)
  const Partials pa = pq_of(c), pb = pr_of(c), pt = pt_of(c);
  const Partials prr{pb.a, nullptr}, prz{pb.b, nullptr};
  const double *rz_part = poly ? prz.a : prr.a;
  const int gz = poly ? gpq : vec_grid4(n);
  // default reliable-update factor: 0.01 with the polynomial (5 FP64
  // refreshes instead of 9 on 256^3, 0.148 -> 0.143 s, at most 3 more
  // iterations on the tested systems), 0.1 with Jacobi (0.01 took the
  // constant 128^3 Laplacian from 332 to 447 Jacobi iterations)
  const double delta = o->delta > 0 ? o->delta : (poly ? 0.01 : 0.1);

  MixedState init{};
  init.tol = o->tol;
  init.delta2 = delta * delta;
  init.max_iter = o->max_iter;
  *c->h_mstate = init;
  HIPCHK(c, hipMemcpyAsync(c->d_mstate, c->h_mstate, sizeof(MixedState), hipMemcpyHostToDevice,
                           c->stream));
  if (int rc = sync(c)) return rc;

  const bool profile = o->profile == CGHIP_PROFILE_KERNELS;
  // Iterations per graph launch: the kernels of a batch that follow a stop or
  // a reliable-update request exit at once but still cost a launch each, so
  // the long polynomial iterations of large systems take small batches (256^3:
  // 155.1-155.3 ms at 4-8, 158.3 ms at 32), the short Jacobi ones large batches
  // (271.8 ms at 8, 270.0 ms at 32).
  // With the polynomial an iteration is degree + 3 kernels and a reliable
  // update falls due every few iterations, after which the rest of the
  // batch still launches (each kernel exits at once): the batch holds about
  // 64 kernels, 3-8 iterations (512k rows, degree 16: 9.4 ms at 32 per
  // graph against 7.6 ms for degree 8).  Jacobi iterations are 3 kernels:
  // 8 per graph from 2^20 rows, else 32.
  const int auto_batch = poly ? std::max(3, std::min(8, 64 / (degree + 3)))
                              : (n >= (1 << 20) ? 8 : 32);
  const int batch = profile ? 1 : o->batch > 0 ? o->batch : std::min(auto_batch, o->max_iter);
  if (profile && ensure_events(c, 7)) return CGHIP_ERR_HIP;
  hipEvent_t *ev = profile ? c->events.data() : nullptr;

  auto read_state = [&]() -> int {
    HIPCHK(c, hipMemcpyAsync(c->h_mstate, c->d_mstate, sizeof(MixedState), hipMemcpyDeviceToHost,
                             c->stream));
    return sync(c);
  };
  // x64 += s o y32 (unless this is the start), r = k - A x64 in FP64, and
  // the inner iteration restarts from r32 = (float)(s o r).
  auto refresh = [&](bool first) -> int {
    if (ev) HIPCHK(c, hipEventRecord(ev[0], c->stream));
    if (!first) HIPCHK(c, mp_accumulate(n, x64, y32, scale, c->stream));
    HIPCHK(c, matvec_crs(crs_of(c), x64, q64, c->stream));
    HIPCHK(c, mp_residual(n, k64, q64, r64, pt, c->stream));
    if (poly) {
      HIPCHK(c, mp_restart(n, r64, scale, r32, y32, p32, false, prr, c->stream));
      HIPCHK(c, mp_poly_apply(A32, r32, c->d_poly, degree, w, first ? p32 : nullptr, prz, nullptr,
                              c->stream, c->poly_fp16));
    } else {
      HIPCHK(c, mp_restart(n, r64, scale, r32, y32, p32, first, prr, c->stream));
    }
    HIPCHK(c, mp_restart_fold(n, pt, prr, rz_part, gz, first, c->d_mstate, h, c->stream));
    if (ev) HIPCHK(c, hipEventRecord(ev[1], c->stream));
    if (int rc = read_state()) return rc;
    if (ev) {
      info->ms_refresh += elapsed(ev[0], ev[1]);
      info->n_refresh++;
    }
    return CGHIP_OK;
  };
  auto enqueue = [&](hipEvent_t *e) -> int {
    if (e) HIPCHK(c, hipEventRecord(e[0], c->stream));
    HIPCHK(c, mp_matvec_dot(A32, p32, q32, pa, c->d_mstate, c->stream));
    if (e) HIPCHK(c, hipEventRecord(e[1], c->stream));
    if (poly) {
      HIPCHK(c, mp_update(n, gpq, r32, q32, pa, prr, c->d_mstate, h, c->stream));
      if (e) HIPCHK(c, hipEventRecord(e[2], c->stream));
      HIPCHK(c, mp_poly_apply(A32, r32, c->d_poly, degree, w, nullptr, prz, c->d_mstate,
                              c->stream, c->poly_fp16));
      if (e) HIPCHK(c, hipEventRecord(e[3], c->stream));
      HIPCHK(c, mp_xpay_poly(n, y32, p32, w[0], prr, prz, gz, c->d_mstate, h, c->stream));
    } else {
      HIPCHK(c, mp_update(n, gpq, r32, q32, pa, prr, c->d_mstate, h, c->stream));
      if (e) HIPCHK(c, hipEventRecord(e[2], c->stream));
      if (e) HIPCHK(c, hipEventRecord(e[3], c->stream));
      HIPCHK(c, mp_xpay(n, y32, p32, r32, prr, c->d_mstate, h, c->stream));
    }
    if (e) HIPCHK(c, hipEventRecord(e[4], c->stream));
    return CGHIP_OK;
  };

  if (!profile && (!c->graph_mp || c->graph_mp_batch != batch || c->graph_mp_precond != o->precond ||
                   c->graph_mp_degree != degree || c->graph_mp_half != (int)c->poly_fp16)) {
    if (c->graph_mp) (void)hipGraphExecDestroy(c->graph_mp);
    c->graph_mp = nullptr;
    hipGraph_t g = nullptr;
    HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    int rc = CGHIP_OK;
    for (int i = 0; i < batch && !rc; i++) rc = enqueue(nullptr);
    hipError_t e = hipStreamEndCapture(c->stream, &g);
    if (rc) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    if (e != hipSuccess) return fail(c, CGHIP_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
    e = hipGraphInstantiate(&c->graph_mp, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
      c->graph_mp = nullptr;
      return fail(c, CGHIP_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
    }
    c->graph_mp_batch = batch;
    c->graph_mp_precond = o->precond;
    c->graph_mp_degree = degree;
    c->graph_mp_half = (int)c->poly_fp16;
  }

  const auto t0 = std::chrono::steady_clock::now();
  if (int rc = refresh(true)) return rc;
  // Each pass either runs a batch, which advances the iteration count by
  // at least one (a reliable update can be requested after the batch's
  // first iteration) until max_iter stops it, or a reliable update (at most
  // one per batch), so the loop ends; the bound below, in iterations rather
  // than batches, only guards against a broken state.
  const int64_t max_passes = 2 * ((int64_t)o->max_iter + 3) + 64;
  for (int64_t pass = 0;; pass++) {
    if (pass >= max_passes)
      return fail(c, CGHIP_ERR_STATE, "mixed solver did not stop after %lld passes",
                  (long long)pass);
    const MixedState &s = *c->h_mstate;
    if (s.next.done == 1) break;  // true residual <= tol
    if (s.next.done == 2) {  // reliable update requested by the inner loop
      const double prev = s.true_res;
      if (int rc = refresh(false)) return rc;
      // A phase that cut the true residual by more than 1 / kResetRatio
      // (e.g. one step down to FP32 resolution) leaves a search direction
      // with no accurate information about the new residual: p = z.
      if (c->h_mstate->true_res < kResetRatio * prev) {
        HIPCHK(c, hipMemcpyAsync(p32, poly ? w[0] : r32, sizeof(float) * (size_t)n,
                                 hipMemcpyDeviceToDevice, c->stream));
        info->restarts++;
      }
      continue;
    }
    if (s.cur.done == 1 || s.next.iteration >= o->max_iter) {
      if (int rc = refresh(false)) return rc;  // fold y32 into x64, final residual
      break;
    }
    if (profile) {
      if (int rc = enqueue(ev + 2)) return rc;
      if (int rc = read_state()) return rc;
      info->ms_matvec += elapsed(ev[2], ev[3]);
      info->ms_update += elapsed(ev[3], ev[4]);
      info->ms_xpay += elapsed(ev[5], ev[6]);
      info->n_matvec++;
      info->n_update++;
      info->n_xpay++;
      if (poly) {
        info->ms_poly += elapsed(ev[4], ev[5]);
        info->n_poly += degree;
      }
    } else {
      HIPCHK(c, hipGraphLaunch(c->graph_mp, c->stream));
      if (int rc = read_state()) return rc;
    }
  }
  const auto t1 = std::chrono::steady_clock::now();
  const MixedState &s = *c->h_mstate;
  c->last_iterations = s.next.iteration;
  info->iterations = s.next.iteration;
  info->updates = s.updates;
  info->true_residual = s.true_res;
  info->inner_residual = s.next.residual;
  info->wall_s = std::chrono::duration<double>(t1 - t0).count();
  return CGHIP_OK;
}

int cghip_history(cghip_ctx *c, int which, double *out, int cap, int *written) {
  if (!c) return CGHIP_ERR_ARG;
  if (which < 0 || which > 4 || !out || cap < 0)
    return fail(c, CGHIP_ERR_ARG, "invalid history request");
  if (!c->d_hist) return fail(c, CGHIP_ERR_STATE, "no solve has run");
  const int64_t have = (which <= 1) ? (int64_t)c->last_iterations + 1 : c->last_iterations;
  const int cnt = (int)std::min<int64_t>(std::min<int64_t>(have, cap), c->hist_cap);
  if (cnt > 0)
    HIPCHK(c, hipMemcpy(out, c->d_hist + (size_t)which * c->hist_cap, sizeof(double) * (size_t)cnt,
                        hipMemcpyDeviceToHost));
  if (written) *written = cnt;
  return CGHIP_OK;
}

int cghip_fill_vector(cghip_ctx *c, int id, double value) {
  if (!c) return CGHIP_ERR_ARG;
  if (!valid_vec(id)) return fail(c, CGHIP_ERR_ARG, "invalid vector id %d", id);
  if (c->n <= 0) return fail(c, CGHIP_ERR_STATE, "no size set");
  HIPCHK(c, fill(c->n, value, c->vec[id], c->stream));
  return sync(c);
}

int cghip_true_residual(cghip_ctx *c, int x, int b, double *norm_r, double *norm_b) {
  if (int rc = check_ready(c)) return rc;
  if (!valid_vec(x) || !valid_vec(b) || x == CGHIP_VEC_Q || b == CGHIP_VEC_Q || !norm_r || !norm_b)
    return fail(c, CGHIP_ERR_ARG, "invalid arguments (x and b must not be Q)");
  HIPCHK(c, launch_matvec(c, c->vec[x], c->vec[CGHIP_VEC_Q]));
  HIPCHK(c, residual_norms(c->n, c->vec[b], c->vec[CGHIP_VEC_Q], pr_of(c), c->dot_out,
                           c->stream));
  HIPCHK(c, hipMemcpyAsync(c->h_scalar, c->dot_out, 2 * sizeof(double), hipMemcpyDeviceToHost,
                           c->stream));
  if (int rc = sync(c)) return rc;
  *norm_r = std::sqrt(c->h_scalar[0]);
  *norm_b = std::sqrt(c->h_scalar[1]);
  return CGHIP_OK;
}

int cghip_matrix_stats(cghip_ctx *c, double *stats, int cap) {
  if (!c) return CGHIP_ERR_ARG;
  if (!stats || cap < CGHIP_MATRIX_STATS) return fail(c, CGHIP_ERR_ARG, "stats needs %d entries", CGHIP_MATRIX_STATS);
  if (!c->has_crs) return fail(c, CGHIP_ERR_STATE, "no CRS matrix uploaded");
  HIPCHK(c, hipSetDevice(c->device));
  const int G = matrix_stats_grid(c->n);
  double *d_part = nullptr;
  if (int rc = dalloc(c, &d_part, (size_t)G * kStatWidth)) return rc;
  std::vector<double> part((size_t)G * kStatWidth);
  hipError_t e = matrix_stats(crs_of(c), d_part, c->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(part.data(), d_part, sizeof(double) * part.size(), hipMemcpyDeviceToHost,
                       c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  dfree(d_part);
  if (e != hipSuccess) return fail(c, CGHIP_ERR_HIP, "matrix statistics: %s", hipGetErrorString(e));
  // fold the blocks in ascending order
  double acc[kStatWidth];
  for (int k = 0; k < kStatWidth; k++) acc[k] = part[k];
  for (int g = 1; g < G; g++) {
    const double *o = part.data() + (size_t)g * kStatWidth;
    for (int q = 0; q < kStatQuantities; q++) {
      for (int k = 0; k < 3; k++) acc[5 * q + k] += o[5 * q + k];
      acc[5 * q + 3] = std::min(acc[5 * q + 3], o[5 * q + 3]);
      acc[5 * q + 4] = std::max(acc[5 * q + 4], o[5 * q + 4]);
    }
    for (int k = 5 * kStatQuantities; k < kStatWidth; k++) acc[k] += o[k];
  }
  const double bad = acc[kStatWidth - 1];
  if (bad > 0)
    return fail(c, CGHIP_ERR_ARG, "%.0f rows have no positive diagonal entry", bad);
  for (int q = 0; q < kStatQuantities; q++) {
    const double n = acc[5 * q], s = acc[5 * q + 1], s2 = acc[5 * q + 2];
    const double mean = n > 0 ? s / n : 0.0;
    const double var = n > 0 ? std::max(s2 / n - mean * mean, 0.0) : 0.0;
    stats[4 * q + 0] = mean;
    stats[4 * q + 1] = std::sqrt(var);
    stats[4 * q + 2] = n > 0 ? acc[5 * q + 3] : 0.0;
    stats[4 * q + 3] = n > 0 ? acc[5 * q + 4] : 0.0;
  }
  for (int b = 0; b < kStatHistBins; b++)
    stats[4 * kStatQuantities + b] = acc[5 * kStatQuantities + b] / (double)c->n;
  return CGHIP_OK;
}

int cghip_mixed_iteration_bytes(const cghip_ctx *c, int precond, double *bytes) {
  if (!c || !bytes) return CGHIP_ERR_ARG;
  auto *cc = const_cast<cghip_ctx *>(c);
  if (!c->has_fp32) return fail(cc, CGHIP_ERR_STATE, "call cghip_prepare_fp32 first");
  const double n = c->n;
  if (precond == CGHIP_PRECOND_NONE) {
    // FP32 CRS copy of A: values, indices, row pointer; x, q; r update (r,
    // q, r); y and p update (y, p, r, y, p)
    *bytes = 8.0 * c->crs_nnz + 4.0 * (n + 1) + 8.0 * n + 12.0 * n + 20.0 * n;
    return CGHIP_OK;
  }
  if (precond != CGHIP_PRECOND_JACOBI && precond != CGHIP_PRECOND_POLY)
    return fail(cc, CGHIP_ERR_ARG, "unknown preconditioner %d", precond);
  if (!c->has_scaled) return fail(cc, CGHIP_ERR_STATE, "no scaled operator");
  if (precond == CGHIP_PRECOND_POLY && !c->has_poly)
    return fail(cc, CGHIP_ERR_STATE, "call cghip_set_poly first");
  const double nnz = c->nnz_nd, d = c->dia_ndiag;
  // operator stream of one SpMV: FP32 values in the CG SpMV (DIA: one
  // stream per diagonal, no mask); the Clenshaw steps read DIA's FP16 row
  // records, or with symmetric storage its upper diagonals (2 bytes per
  // stored pair; the lower values are those of earlier rows)
  const double op32 = c->dia_nd ? 4.0 * d * n : c->ell_nd ? 8.0 * nnz + 4.0 * n
                                                          : 8.0 * nnz + 4.0 * (n + 1);
  const double op16 = !(c->dia_nd && c->poly_fp16) ? op32
                      : c->dia_up16                ? 1.0 * d * n
                                                   : 4.0 * half_words(c->dia_ndiag) * n;
  const double spmv = op32 + 8.0 * n;
  if (precond == CGHIP_PRECOND_JACOBI) {
    *bytes = spmv + 12.0 * n + 20.0 * n;  // + r update (r, q; r) + y, p update (y, p, r; y, p)
    return CGHIP_OK;
  }
  const int k = c->poly_degree;
  // r update (r, q; r).  Clenshaw steps: the first reads r (gathered, as
  // B_k = c_k r) and writes B_{k-1}; the second reads B_{k-1}, r (also as
  // B_k) and writes B_{k-2}; every later one reads B_{j+1}, r, B_{j+2} and
  // writes B_j.  y, p update reads y, p, z and writes y, p.
  const double steps = k * op16 + 8.0 * n + (k >= 2 ? 12.0 * n : 0.0) +
                       (k > 2 ? (k - 2) * 16.0 * n : 0.0);
  *bytes = spmv + 12.0 * n + steps + 20.0 * n;
  return CGHIP_OK;
}

int cghip_time_poly_step(cghip_ctx *c, int reps, double *avg_ms) {
  if (!c) return CGHIP_ERR_ARG;
  if (reps <= 0 || !avg_ms) return fail(c, CGHIP_ERR_ARG, "invalid arguments");
  if (!c->has_poly || !c->has_scaled) return fail(c, CGHIP_ERR_STATE, "call cghip_set_poly first");
  HIPCHK(c, hipSetDevice(c->device));
  if (int rc = ensure_events(c, 2)) return rc;
  const Op32 A = op32_of(c, true);
  float *const w[3] = {c->w32[0], c->w32[1], c->w32[2]};
  const Partials prz{pr_of(c).b, nullptr};
  // defined inputs: r32 = 1, B_k = c_k r32 is rebuilt by each application
  HIPCHK(c, fill32(c->n, 1.0f, c->v32[1], c->stream));
  for (auto &v : c->w32) HIPCHK(c, fill32(c->n, 0.0f, v, c->stream));
  HIPCHK(c, mp_poly_apply(A, c->v32[1], c->d_poly, c->poly_degree, w, nullptr, prz, nullptr,
                          c->stream, c->poly_fp16));  // warm-up
  HIPCHK(c, hipEventRecord(c->events[0], c->stream));
  for (int i = 0; i < reps; i++)
    HIPCHK(c, mp_poly_apply(A, c->v32[1], c->d_poly, c->poly_degree, w, nullptr, prz, nullptr,
                            c->stream, c->poly_fp16));
  HIPCHK(c, hipEventRecord(c->events[1], c->stream));
  HIPCHK(c, hipEventSynchronize(c->events[1]));
  *avg_ms = elapsed(c->events[0], c->events[1]) / ((double)reps * c->poly_degree);
  return CGHIP_OK;
}

int cghip_time_matvec(cghip_ctx *c, int x, int y, int reps, double *avg_ms) {
  if (int rc = check_ready(c)) return rc;
  if (!valid_vec(x) || !valid_vec(y) || x == y || reps <= 0 || !avg_ms)
    return fail(c, CGHIP_ERR_ARG, "invalid arguments");
  if (int rc = ensure_events(c, 2)) return rc;
  HIPCHK(c, launch_matvec(c, c->vec[x], c->vec[y]));  // warm-up
  HIPCHK(c, hipEventRecord(c->events[0], c->stream));
  for (int i = 0; i < reps; i++) HIPCHK(c, launch_matvec(c, c->vec[x], c->vec[y]));
  HIPCHK(c, hipEventRecord(c->events[1], c->stream));
  HIPCHK(c, hipEventSynchronize(c->events[1]));
  *avg_ms = elapsed(c->events[0], c->events[1]) / reps;
  return CGHIP_OK;
}

int cghip_time_hook(cghip_ctx *c, int hook, int reps, double *avg_ms) {
  if (!c) return CGHIP_ERR_ARG;
  if (hook < CGHIP_HOOK_MATVEC || hook > CGHIP_HOOK_CPY || reps <= 0 || !avg_ms)
    return fail(c, CGHIP_ERR_ARG, "invalid arguments");
  if (c->n <= 0) return fail(c, CGHIP_ERR_STATE, "no size set");
  if (hook == CGHIP_HOOK_MATVEC)
    if (int rc = check_ready(c)) return rc;
  if (hook == CGHIP_HOOK_JACOBI && !c->has_jac)
    return fail(c, CGHIP_ERR_STATE, "no Jacobi preconditioner uploaded");
  HIPCHK(c, hipSetDevice(c->device));
  if (int rc = ensure_events(c, 2)) return rc;
  double *P = c->vec[CGHIP_VEC_P], *Q = c->vec[CGHIP_VEC_Q], *R = c->vec[CGHIP_VEC_R],
         *Z = c->vec[CGHIP_VEC_Z];
  const size_t bytes = sizeof(double) * (size_t)c->n;
  auto launch = [&]() -> hipError_t {
    switch (hook) {
      case CGHIP_HOOK_MATVEC: return launch_matvec(c, P, Q);
      case CGHIP_HOOK_AXPY: return axpy(c->n, 1e-3, P, Q, c->stream);
      case CGHIP_HOOK_XPAY: return xpay(c->n, P, 0.5, R, c->stream);
      case CGHIP_HOOK_DOT:
        return c->dot_mode == CGHIP_DOT_SERIAL
                   ? dot_serial(c->n, P, R, c->dot_out, c->stream)
                   : dot_tree(c->n, P, R, pr_of(c), c->dot_out, c->stream);
      case CGHIP_HOOK_JACOBI: return jacobi(c->n, c->jac, R, Z, c->stream);
      default: return hipMemcpyAsync(Q, P, bytes, hipMemcpyDeviceToDevice, c->stream);
    }
  };
  HIPCHK(c, launch());  // warm-up
  HIPCHK(c, hipEventRecord(c->events[0], c->stream));
  for (int i = 0; i < reps; i++) HIPCHK(c, launch());
  HIPCHK(c, hipEventRecord(c->events[1], c->stream));
  HIPCHK(c, hipEventSynchronize(c->events[1]));
  *avg_ms = elapsed(c->events[0], c->events[1]) / reps;
  return CGHIP_OK;
}

}  // extern "C"
