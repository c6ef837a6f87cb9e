// Allocation-free dot product (cghip_dot_noalloc, include/cghip.h).
//
// For callers that cannot allocate device memory: code under HIP graph
// capture, tight serving loops, harnesses that forbid hipMalloc.  The
// kernels are those of the tree-ordered dot of cghip_vector_dot (per-block
// partials, then a one-block fold), so the result is the same bit for bit;
// only the partials live elsewhere: in `scratch`, a file-scope device array
// of kNoAllocPartials doubles that every GPU the library runs on holds until
// the process ends.  Every call on one GPU shares that GPU's copy: a call
// writes all the partials it folds, so nothing has to start at zero, but
// calls that overlap on one GPU (two streams, two threads) race on it.
//
// The array's name is a host-side stand-in in host code, not a device
// address: its address on each device comes from hipGetSymbolAddress, looked
// up once per device with that device current (by cghip_create, or by the
// first call of the entry point on that device) and kept in `g_scratch`.
// Kernels get that pointer, never the name.

#include "../../include/cghip.h"
#include "cghip_kernels.h"

#include <cstdarg>
#include <cstdio>
#include <mutex>
#include <string>

namespace cghip {

// Partials of one tree dot: vec_grid(n) <= kVecGMax blocks.
constexpr int kNoAllocPartials = kVecGMax;
constexpr int kNoAllocDevices = 64;

namespace {

__device__ __align__(256) double scratch[kNoAllocPartials];

std::mutex g_mu;
double *g_scratch[kNoAllocDevices] = {};

// The scratch array's address on `device` (which must be current).
hipError_t scratch_on(int device, double **out) {
  if (device < 0 || device >= kNoAllocDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_scratch[device]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch[device] = static_cast<double *>(p);
  }
  *out = g_scratch[device];
  return hipSuccess;
}

}  // namespace

// Looks the array up on `device` (current); called by cghip_create.
hipError_t noalloc_prepare(int device) {
  double *p = nullptr;
  return scratch_on(device, &p);
}

}  // namespace cghip

using namespace cghip;

namespace {

int noalloc_fail(int code, const char *fmt, ...) {
  char buf[256];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  set_global_error(buf);
  return code;
}

}  // namespace

extern "C" int cghip_dot_noalloc(const double *a, const double *b, int n, double *out,
                                 void *stream) {
  if (n < 0) return noalloc_fail(CGHIP_ERR_ARG, "cghip_dot_noalloc: n = %d is negative", n);
  if (!out || (n > 0 && (!a || !b)))
    return noalloc_fail(CGHIP_ERR_ARG, "cghip_dot_noalloc: NULL vector or output pointer");
  const int G = vec_grid(n);
  if (G > kNoAllocPartials)  // vec_grid caps at kVecGMax: a guard, never taken
    return noalloc_fail(CGHIP_ERR_ARG, "cghip_dot_noalloc: %d partials exceed the %d of the scratch",
                        G, kNoAllocPartials);
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  double *part = nullptr;
  if (e == hipSuccess) e = scratch_on(dev, &part);
  if (e == hipSuccess)
    e = dot_tree(n, a, b, Partials{part, nullptr}, out, static_cast<hipStream_t>(stream));
  if (e != hipSuccess)
    return noalloc_fail(CGHIP_ERR_HIP, "cghip_dot_noalloc: %s", hipGetErrorString(e));
  return CGHIP_OK;
}
