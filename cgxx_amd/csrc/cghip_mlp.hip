// C ABI of the FP32 network (declared in include/cghip_mlp.h).
//
// The handle owns one stream and the device buffers: parameters, Adam
// moments, the inputs/targets of up to max_rows rows, the activations, the
// partial gradients of the backward blocks and the per-step losses.

#include "../../include/cghip_mlp.h"
#include "mlp.h"

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <string>
#include <vector>

using namespace cghip;

namespace {

constexpr int kBackwardBlocks = 128;  // partial-gradient slices

bool valid_dims(int in, int h1, int h2, int out) {
  auto ok = [](int w) { return w >= 1 && w <= kMlpMaxWidth; };
  return ok(in) && ok(h1) && ok(h2) && ok(out);
}

std::string g_err = "no error";

}  // namespace

struct cghip_mlp {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  MlpDims d{};
  int total = 0, max_rows = 0, step = 0;
  float *P = nullptr, *M = nullptr, *V = nullptr, *grad = nullptr;
  float *X = nullptr, *T = nullptr, *A1 = nullptr, *A2 = nullptr, *Y = nullptr, *Gp = nullptr;
  double *Lp = nullptr, *loss = nullptr;
  int loss_cap = 0;
};

namespace {

int fail(cghip_mlp *m, int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  (m ? m->err : g_err) = buf;
  return code;
}

#define MLPCHK(m, call)                                                                  \
  do {                                                                                   \
    hipError_t e_ = (call);                                                              \
    if (e_ != hipSuccess) return fail((m), 2, "%s failed: %s", #call, hipGetErrorString(e_)); \
  } while (0)

template <typename T>
int alloc(cghip_mlp *m, T **p, size_t count) {
  *p = nullptr;
  hipError_t e = hipMalloc((void **)p, std::max<size_t>(count, 1) * sizeof(T));
  if (e != hipSuccess) {
    *p = nullptr;
    return fail(m, 4, "hipMalloc of %zu bytes failed: %s", count * sizeof(T), hipGetErrorString(e));
  }
  return 0;
}

template <typename T>
void release(T *&p) {
  if (p) (void)hipFree((void *)p);
  p = nullptr;
}

int check_rows(cghip_mlp *m, int rows) {
  if (rows < 1 || rows > m->max_rows)
    return fail(m, 1, "rows = %d outside [1, %d]", rows, m->max_rows);
  return 0;
}

int sync(cghip_mlp *m) {
  MLPCHK(m, hipStreamSynchronize(m->stream));
  return 0;
}

// Forward with activations kept, then the backward pass into Gp / Lp.
int forward_backward(cghip_mlp *m, int rows, int G) {
  MLPCHK(m, mlp_forward(m->d, m->P, m->X, rows, m->A1, m->A2, m->Y, m->stream));
  MLPCHK(m, mlp_backward(m->d, m->P, m->X, m->A1, m->A2, m->Y, m->T, rows, G, m->Gp, m->Lp,
                         m->stream));
  return 0;
}

int upload_rows(cghip_mlp *m, const float *X, const float *T, int rows) {
  MLPCHK(m, hipMemcpyAsync(m->X, X, sizeof(float) * (size_t)rows * m->d.in, hipMemcpyHostToDevice,
                           m->stream));
  if (T)
    MLPCHK(m, hipMemcpyAsync(m->T, T, sizeof(float) * (size_t)rows * m->d.out,
                             hipMemcpyHostToDevice, m->stream));
  return 0;
}

}  // namespace

extern "C" {

int cghip_mlp_param_count(int in, int h1, int h2, int out) {
  if (!valid_dims(in, h1, h2, out)) return 0;
  return mlp_param_count(MlpDims{in, h1, h2, out});
}

const char *cghip_mlp_last_error(const cghip_mlp *m) { return m ? m->err.c_str() : g_err.c_str(); }

int cghip_mlp_destroy(cghip_mlp *m) {
  if (!m) return 0;
  (void)hipSetDevice(m->device);
  if (m->stream) (void)hipStreamSynchronize(m->stream);
  release(m->P), release(m->M), release(m->V), release(m->grad);
  release(m->X), release(m->T), release(m->A1), release(m->A2), release(m->Y), release(m->Gp);
  release(m->Lp), release(m->loss);
  if (m->stream) (void)hipStreamDestroy(m->stream);
  delete m;
  return 0;
}

int cghip_mlp_create(int device, int in, int h1, int h2, int out, int max_rows, cghip_mlp **res) {
  if (!res) return fail(nullptr, 1, "out is NULL");
  *res = nullptr;
  if (!valid_dims(in, h1, h2, out))
    return fail(nullptr, 1, "layer widths must be in [1, %d]", kMlpMaxWidth);
  if (max_rows < 1) return fail(nullptr, 1, "max_rows must be positive");
  int count = 0;
  hipError_t e = hipGetDeviceCount(&count);
  if (e != hipSuccess || count <= 0)
    return fail(nullptr, 2, "no HIP device available (%s)", hipGetErrorString(e));
  if (device < 0 || device >= count) return fail(nullptr, 1, "device %d out of range", device);
  if ((e = hipSetDevice(device)) != hipSuccess)
    return fail(nullptr, 2, "hipSetDevice: %s", hipGetErrorString(e));
  cghip_mlp *m = new cghip_mlp;
  m->device = device;
  m->d = MlpDims{in, h1, h2, out};
  m->total = mlp_param_count(m->d);
  m->max_rows = max_rows;
  int rc = 0;
  if ((e = hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking)) != hipSuccess)
    rc = fail(m, 2, "hipStreamCreate: %s", hipGetErrorString(e));
  const size_t R = (size_t)max_rows, P = (size_t)m->total;
  if (!rc) rc = alloc(m, &m->P, P);
  if (!rc) rc = alloc(m, &m->M, P);
  if (!rc) rc = alloc(m, &m->V, P);
  if (!rc) rc = alloc(m, &m->grad, P);
  if (!rc) rc = alloc(m, &m->X, R * in);
  if (!rc) rc = alloc(m, &m->T, R * out);
  if (!rc) rc = alloc(m, &m->A1, R * h1);
  if (!rc) rc = alloc(m, &m->A2, R * h2);
  if (!rc) rc = alloc(m, &m->Y, R * out);
  if (!rc) rc = alloc(m, &m->Gp, (size_t)kBackwardBlocks * P);
  if (!rc) rc = alloc(m, &m->Lp, kBackwardBlocks);
  if (!rc && (e = hipMemsetAsync(m->P, 0, sizeof(float) * P, m->stream)) != hipSuccess)
    rc = fail(m, 2, "hipMemset: %s", hipGetErrorString(e));
  if (!rc && (e = hipMemsetAsync(m->M, 0, sizeof(float) * P, m->stream)) != hipSuccess)
    rc = fail(m, 2, "hipMemset: %s", hipGetErrorString(e));
  if (!rc && (e = hipMemsetAsync(m->V, 0, sizeof(float) * P, m->stream)) != hipSuccess)
    rc = fail(m, 2, "hipMemset: %s", hipGetErrorString(e));
  if (!rc) rc = sync(m);
  if (rc) {
    g_err = m->err;
    cghip_mlp_destroy(m);
    return rc;
  }
  *res = m;
  return 0;
}

int cghip_mlp_set_params(cghip_mlp *m, const float *params) {
  if (!m) return 1;
  if (!params) return fail(m, 1, "params is NULL");
  for (int k = 0; k < m->total; k++)
    if (!std::isfinite(params[k])) return fail(m, 1, "parameter %d is not finite", k);
  MLPCHK(m, hipSetDevice(m->device));
  MLPCHK(m, hipMemcpyAsync(m->P, params, sizeof(float) * m->total, hipMemcpyHostToDevice, m->stream));
  MLPCHK(m, hipMemsetAsync(m->M, 0, sizeof(float) * m->total, m->stream));
  MLPCHK(m, hipMemsetAsync(m->V, 0, sizeof(float) * m->total, m->stream));
  m->step = 0;
  return sync(m);
}

int cghip_mlp_get_params(cghip_mlp *m, float *params) {
  if (!m) return 1;
  if (!params) return fail(m, 1, "params is NULL");
  MLPCHK(m, hipSetDevice(m->device));
  MLPCHK(m, hipMemcpyAsync(params, m->P, sizeof(float) * m->total, hipMemcpyDeviceToHost, m->stream));
  return sync(m);
}

int cghip_mlp_forward(cghip_mlp *m, const float *X, int rows, float *Y) {
  if (!m) return 1;
  if (!X || !Y) return fail(m, 1, "X/Y is NULL");
  if (int rc = check_rows(m, rows)) return rc;
  MLPCHK(m, hipSetDevice(m->device));
  if (int rc = upload_rows(m, X, nullptr, rows)) return rc;
  MLPCHK(m, mlp_forward(m->d, m->P, m->X, rows, nullptr, nullptr, m->Y, m->stream));
  MLPCHK(m, hipMemcpyAsync(Y, m->Y, sizeof(float) * (size_t)rows * m->d.out, hipMemcpyDeviceToHost,
                           m->stream));
  return sync(m);
}

int cghip_mlp_gradient(cghip_mlp *m, const float *X, const float *T, int rows, float *grad,
                       double *loss) {
  if (!m) return 1;
  if (!X || !T || !grad || !loss) return fail(m, 1, "NULL argument");
  if (int rc = check_rows(m, rows)) return rc;
  MLPCHK(m, hipSetDevice(m->device));
  if (int rc = upload_rows(m, X, T, rows)) return rc;
  const int G = std::min(rows, kBackwardBlocks);
  if (int rc = forward_backward(m, rows, G)) return rc;
  AdamStep a{};
  a.apply = 0;
  MLPCHK(m, mlp_adam(m->total, G, m->Gp, m->grad, m->P, m->M, m->V, a, m->stream));
  if (m->loss_cap < 1) {
    release(m->loss);
    m->loss_cap = 0;
    if (int rc = alloc(m, &m->loss, 1)) return rc;
    m->loss_cap = 1;
  }
  MLPCHK(m, mlp_loss(G, m->Lp, (double)rows * m->d.out, m->loss, m->stream));
  MLPCHK(m, hipMemcpyAsync(grad, m->grad, sizeof(float) * m->total, hipMemcpyDeviceToHost, m->stream));
  MLPCHK(m, hipMemcpyAsync(loss, m->loss, sizeof(double), hipMemcpyDeviceToHost, m->stream));
  return sync(m);
}

int cghip_mlp_train(cghip_mlp *m, const float *X, const float *T, int rows, int steps,
                    const cghip_adam_opts *o, double *loss_history) {
  if (!m) return 1;
  if (!X || !T || !o) return fail(m, 1, "NULL argument");
  if (int rc = check_rows(m, rows)) return rc;
  if (steps < 0) return fail(m, 1, "steps must be >= 0");
  if (!(o->lr > 0) || !(o->beta1 >= 0 && o->beta1 < 1) || !(o->beta2 >= 0 && o->beta2 < 1) ||
      !(o->eps > 0))
    return fail(m, 1, "invalid Adam options (lr > 0, 0 <= beta < 1, eps > 0)");
  MLPCHK(m, hipSetDevice(m->device));
  if (int rc = upload_rows(m, X, T, rows)) return rc;
  if (m->loss_cap < std::max(steps, 1)) {
    release(m->loss);
    m->loss_cap = 0;
    if (int rc = alloc(m, &m->loss, (size_t)std::max(steps, 1))) return rc;
    m->loss_cap = std::max(steps, 1);
  }
  const int G = std::min(rows, kBackwardBlocks);
  for (int s = 0; s < steps; s++) {
    if (int rc = forward_backward(m, rows, G)) return rc;
    MLPCHK(m, mlp_loss(G, m->Lp, (double)rows * m->d.out, m->loss + s, m->stream));
    const int t = ++m->step;
    AdamStep a;
    a.lr = o->lr;
    a.beta1 = o->beta1;
    a.beta2 = o->beta2;
    a.eps = o->eps;
    a.corr1 = (float)(1.0 / (1.0 - std::pow((double)o->beta1, t)));
    a.corr2 = (float)(1.0 / (1.0 - std::pow((double)o->beta2, t)));
    a.apply = 1;
    MLPCHK(m, mlp_adam(m->total, G, m->Gp, nullptr, m->P, m->M, m->V, a, m->stream));
  }
  if (loss_history && steps > 0)
    MLPCHK(m, hipMemcpyAsync(loss_history, m->loss, sizeof(double) * steps, hipMemcpyDeviceToHost,
                             m->stream));
  return sync(m);
}

}  // extern "C"
