// Mixed-precision PCG kernels for gfx950: an FP32 inner iteration with FP64
// reliable updates (Sleijpen & van der Vorst, "Reliable updated residuals in
// hybrid Bi-CG methods", Computing 56, 1996; the mixed-precision form of
// Clark et al., "Solving lattice QCD systems of equations using mixed
// precision solvers on GPUs", Comput. Phys. Commun. 181, 2010).
//
// The inner iteration is the fused three-kernel PCG of kernels.hip on FP32
// data: the FP32 copy of the CRS values (8 bytes per nonzero with the int32
// index instead of 12), FP32 vectors (half the vector traffic).  Every dot
// product is accumulated in FP64 from the products (double)a * (double)b,
// which are exact for float operands, so only the summation rounds, in a
// fixed tree order.  alpha and beta are computed in FP64 and rounded to
// FP32 for the vector updates.  The solution and the true residual live in
// FP64: at each reliable update x64 += x32 and r = k - A x64 is recomputed
// with the FP64 matrix; the search direction p32 is kept, so no Krylov
// information is lost as it would be by restarting.  The stop test is the
// FP64 true relative residual.
//
// Vector kernels on FP32 data work on element quads (16-byte loads).

#include "device_common.h"

#pragma clang fp contract(off)

namespace cghip {

namespace {

using namespace dev;

constexpr int kQuadUnroll = 2;  // quads in flight per thread per vector

__global__ __launch_bounds__(kBlock) void k_convert(int64_t n, const double *__restrict__ x,
                                                    float *__restrict__ y) {
  const int64_t T = (int64_t)gridDim.x * kBlock;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += T) y[i] = (float)x[i];
}

__global__ __launch_bounds__(kBlock) void k_mp_accum(int n, double *__restrict__ x64,
                                                     const float *__restrict__ x32) {
  const int64_t T = (int64_t)gridDim.x * kBlock;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += T)
    x64[i] = x64[i] + (double)x32[i];
}

// r64 = k - q64 and the partials of ||r64||^2 (pairs order).
__global__ __launch_bounds__(kBlock) void k_mp_residual(int n, const double *__restrict__ k,
                                                        const double *__restrict__ q,
                                                        double *__restrict__ r, Partials pa) {
  __shared__ double lds[4];
  Pairs P(n);
  double s = 0.0;
  for (; P.q < P.npair; P.q += P.T) {
    for (int64_t e = 2 * P.q; e < 2 * P.q + 2 && e < n; e++) {
      const double ri = k[e] - q[e];
      r[e] = ri;
      s += ri * ri;
    }
  }
  const double bp = block_fold(s, lds);
  if (threadIdx.x == 0) pa.a[blockIdx.x] = bp;
}

// r32 = (float)r64, x32 = 0, z = C32 r32, p32 = z when first; partials of
// <r32,r32> and <r32,z> (quads order).
template <bool JAC>
__global__ __launch_bounds__(kBlock) void k_mp_restart(int n, const double *__restrict__ r64,
                                                       float *__restrict__ r32,
                                                       float *__restrict__ x32,
                                                       float *__restrict__ p32,
                                                       const float *__restrict__ C32, int first,
                                                       Partials pb) {
  __shared__ double lds[4];
  Quads Q(n);
  double s_rr = 0.0, s_rz = 0.0;
  for (; Q.q < Q.nquad; Q.q += Q.T) {
    for (int64_t e = 4 * Q.q; e < 4 * Q.q + 4 && e < n; e++) {
      const float r = (float)r64[e];
      r32[e] = r;
      x32[e] = 0.0f;
      const float z = JAC ? C32[e] * r : r;
      if (first) p32[e] = z;
      s_rr += (double)r * (double)r;
      s_rz += (double)r * (double)z;
    }
  }
  const double b0 = block_fold(s_rr, lds);
  const double b1 = block_fold(s_rz, lds);
  if (threadIdx.x == 0) {
    pb.a[blockIdx.x] = b0;
    pb.b[blockIdx.x] = b1;
  }
}

// One block: the true residual norm, rho and r2 of the restarted inner
// iteration, the stop test on the true residual.
__global__ __launch_bounds__(kBlock) void k_mp_restart_fold(Partials prt, int gt, Partials prr, int gr,
                                                            const double *rzp, int gz, int first,
                                                            MixedState *st, History h) {
  __shared__ double lds[4];
  const double rt2 = fold_partials(prt.a, gt, lds);
  const double rr = fold_partials(prr.a, gr, lds);
  const double rz = fold_partials(rzp, gz, lds);
  if (threadIdx.x == 0) {
    MixedScalars s = st->next;
    if (first) {
      st->nrm2_0 = sqrt(rt2);
      st->updates = 0;
      s.iteration = 0;
    } else {
      st->updates += 1;
    }
    const double nrm0 = st->nrm2_0;
    st->true_res = sqrt(rt2) / nrm0;
    s.rho = rz;
    s.r2 = rr;
    s.r2_ref = rr;
    s.residual = sqrt(rr) / nrm0;
    s.done = st->true_res <= st->tol ? 1 : 0;
    st->next = s;
    st->cur = s;
    if (first) {
      if (h.r2) h.r2[0] = rr;
      if (h.rho) h.rho[0] = rz;
    }
  }
}

__device__ inline bool mp_begin(MixedState *st) {
  const MixedScalars s = st->next;
  const bool stop = s.done != 0 || s.iteration >= st->max_iter;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    MixedScalars c = s;
    if (stop && c.done == 0) c.done = 1;
    st->cur = c;
  }
  return !stop;
}

// q32 = A32 p32 (rows summed left to right in FP32) and partials of <p,q>.
__global__ __launch_bounds__(kBlock) void k_mp_matvec_dot_crs(Crs<float> A, const float *__restrict__ p,
                                                              float *__restrict__ q, Partials pq,
                                                              MixedState *st) {
  __shared__ float prod[kNnzCap];
  __shared__ double lds[4];
  if (!mp_begin(st)) return;
  double tpart = 0.0;
  for (int rb = blockIdx.x; rb < A.nrb; rb += gridDim.x)
    tpart += crs_rowblock<float, float, true>(A, p, q, rb, prod);
  const double bp = block_fold(tpart, lds);
  if (threadIdx.x == 0) pq.a[blockIdx.x] = bp;
}

template <bool JAC>
__device__ inline void mp_update_elem(float &x, float &r, float p, float q, float c, float a,
                                      float na, double &s_rr, double &s_rz) {
  x = x + a * p;
  r = r + na * q;
  const float z = JAC ? c * r : r;
  s_rr += (double)r * (double)r;
  s_rz += (double)r * (double)z;
}

// alpha = rho / <p,q>; x32 += alpha p32; r32 -= alpha q32; partials of
// <r,r> and <r,z>, z = C32 r32.
template <bool JAC>
__global__ __launch_bounds__(kBlock) void k_mp_update(int n, int gpq, float *__restrict__ x,
                                                      float *__restrict__ r,
                                                      const float *__restrict__ p,
                                                      const float *__restrict__ q,
                                                      const float *__restrict__ C, Partials ppq,
                                                      Partials prr, const MixedState *st,
                                                      History h) {
  __shared__ double lds[4];
  const MixedScalars c = st->cur;
  if (c.done) return;
  const double pq = fold_partials(ppq.a, gpq, lds);
  const double alpha = c.rho / pq;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    if (h.pq) h.pq[c.iteration] = pq;
    if (h.alpha) h.alpha[c.iteration] = alpha;
  }
  const float a = (float)alpha, na = (float)(-alpha);
  Quads Q(n);
  double s_rr = 0.0, s_rz = 0.0;
  for (; Q.q + (kQuadUnroll - 1) * Q.T < Q.full; Q.q += kQuadUnroll * Q.T) {
    float4 xv[kQuadUnroll], rv[kQuadUnroll], pv[kQuadUnroll], qv[kQuadUnroll], cv[kQuadUnroll];
#pragma unroll
    for (int u = 0; u < kQuadUnroll; u++) {
      const int64_t j = Q.q + u * Q.T;
      xv[u] = v4(x)[j];
      rv[u] = v4(r)[j];
      pv[u] = v4(p)[j];
      qv[u] = v4(q)[j];
      if (JAC) cv[u] = v4(C)[j];
    }
#pragma unroll
    for (int u = 0; u < kQuadUnroll; u++) {
      const int64_t j = Q.q + u * Q.T;
      mp_update_elem<JAC>(xv[u].x, rv[u].x, pv[u].x, qv[u].x, JAC ? cv[u].x : 0.f, a, na, s_rr, s_rz);
      mp_update_elem<JAC>(xv[u].y, rv[u].y, pv[u].y, qv[u].y, JAC ? cv[u].y : 0.f, a, na, s_rr, s_rz);
      mp_update_elem<JAC>(xv[u].z, rv[u].z, pv[u].z, qv[u].z, JAC ? cv[u].z : 0.f, a, na, s_rr, s_rz);
      mp_update_elem<JAC>(xv[u].w, rv[u].w, pv[u].w, qv[u].w, JAC ? cv[u].w : 0.f, a, na, s_rr, s_rz);
      v4(x)[j] = xv[u];
      v4(r)[j] = rv[u];
    }
  }
  for (; Q.q < Q.nquad; Q.q += Q.T) {
    for (int64_t e = 4 * Q.q; e < 4 * Q.q + 4 && e < n; e++) {
      float xe = x[e], re = r[e];
      mp_update_elem<JAC>(xe, re, p[e], q[e], JAC ? C[e] : 0.f, a, na, s_rr, s_rz);
      x[e] = xe;
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

// Recursive residual test and reliable-update trigger; beta = rho_new / rho;
// p32 = z + beta p32 (always: a reliable update keeps p).  Writes `next`.
template <bool JAC>
__global__ __launch_bounds__(kBlock) void k_mp_xpay(int n, float *__restrict__ p,
                                                    const float *__restrict__ r,
                                                    const float *__restrict__ C, Partials prr,
                                                    MixedState *st, History h) {
  __shared__ double lds[4];
  const MixedScalars c = st->cur;
  if (c.done) return;
  const int G = gridDim.x;  // the update kernel ran on the same grid
  const double rr = fold_partials(prr.a, G, lds);
  const double rz = fold_partials(prr.b, G, lds);
  const double res = sqrt(rr) / st->nrm2_0;
  const double beta = rz / c.rho;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    const int it = c.iteration;
    MixedScalars s = c;
    s.iteration = it + 1;
    s.rho = rz;
    s.r2 = rr;
    s.residual = res;
    if (res <= st->tol || rr <= st->delta2 * c.r2_ref) s.done = 2;
    st->next = s;
    if (h.r2) h.r2[it + 1] = rr;
    if (h.rho) h.rho[it + 1] = rz;
    if (h.beta) h.beta[it] = beta;
  }
  const float b = (float)beta;
  Quads Q(n);
  for (; Q.q + (kQuadUnroll - 1) * Q.T < Q.full; Q.q += kQuadUnroll * Q.T) {
    float4 pv[kQuadUnroll], rv[kQuadUnroll], cv[kQuadUnroll];
#pragma unroll
    for (int u = 0; u < kQuadUnroll; u++) {
      const int64_t j = Q.q + u * Q.T;
      pv[u] = v4(p)[j];
      rv[u] = v4(r)[j];
      if (JAC) cv[u] = v4(C)[j];
    }
#pragma unroll
    for (int u = 0; u < kQuadUnroll; u++) {
      float4 o;
      o.x = (JAC ? cv[u].x * rv[u].x : rv[u].x) + b * pv[u].x;
      o.y = (JAC ? cv[u].y * rv[u].y : rv[u].y) + b * pv[u].y;
      o.z = (JAC ? cv[u].z * rv[u].z : rv[u].z) + b * pv[u].z;
      o.w = (JAC ? cv[u].w * rv[u].w : rv[u].w) + b * pv[u].w;
      v4(p)[Q.q + u * Q.T] = o;
    }
  }
  for (; Q.q < Q.nquad; Q.q += Q.T) {
    for (int64_t e = 4 * Q.q; e < 4 * Q.q + 4 && e < n; e++) {
      const float z = JAC ? C[e] * r[e] : r[e];
      p[e] = z + b * p[e];
    }
  }
}

// ---------------------------------------------------------------------------
// Polynomial preconditioner (Clenshaw steps fused into FP32 SpMVs).

// x32 += alpha p32; r32 -= alpha q32; B_k = c_k (C32 r32); partials of <r,r>.
__global__ __launch_bounds__(kBlock) void k_mp_update_poly(int n, int gpq, float *__restrict__ x,
                                                           float *__restrict__ r,
                                                           const float *__restrict__ p,
                                                           const float *__restrict__ q,
                                                           const float *__restrict__ C,
                                                           const PolyParams *poly,
                                                           float *__restrict__ bk, Partials ppq,
                                                           Partials prr, const MixedState *st,
                                                           History h) {
  __shared__ double lds[4];
  const MixedScalars c = st->cur;
  if (c.done) return;
  const double pq = fold_partials(ppq.a, gpq, lds);
  const double alpha = c.rho / pq;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    if (h.pq) h.pq[c.iteration] = pq;
    if (h.alpha) h.alpha[c.iteration] = alpha;
  }
  const float a = (float)alpha, na = (float)(-alpha);
  const float ck = poly->c[poly->degree];
  Quads Q(n);
  double s_rr = 0.0;
  for (; Q.q < Q.nquad; Q.q += Q.T) {
    for (int64_t e = 4 * Q.q; e < 4 * Q.q + 4 && e < n; e++) {
      x[e] = x[e] + a * p[e];
      const float re = r[e] + na * q[e];
      r[e] = re;
      bk[e] = ck * (C[e] * re);
      s_rr += (double)re * (double)re;
    }
  }
  const double b0 = block_fold(s_rr, lds);
  if (threadIdx.x == 0) prr.a[blockIdx.x] = b0;
}

// Restart of the polynomial variant: r32 = (float)r64, x32 = 0,
// B_k = c_k (C32 r32), partials of <r,r>.
__global__ __launch_bounds__(kBlock) void k_mp_restart_poly(int n, const double *__restrict__ r64,
                                                            float *__restrict__ r32,
                                                            float *__restrict__ x32,
                                                            const float *__restrict__ C,
                                                            const PolyParams *poly,
                                                            float *__restrict__ bk, Partials prr) {
  __shared__ double lds[4];
  const float ck = poly->c[poly->degree];
  Quads Q(n);
  double s_rr = 0.0;
  for (; Q.q < Q.nquad; Q.q += Q.T) {
    for (int64_t e = 4 * Q.q; e < 4 * Q.q + 4 && e < n; e++) {
      const float re = (float)r64[e];
      r32[e] = re;
      x32[e] = 0.0f;
      bk[e] = ck * (C[e] * re);
      s_rr += (double)re * (double)re;
    }
  }
  const double b0 = block_fold(s_rr, lds);
  if (threadIdx.x == 0) prr.a[blockIdx.x] = b0;
}

// Clenshaw step j with one SpMV t = A B_{j+1}:
//   s = (2/b) (C t) - B_{j+1}
//   j > 0: B_j = (c_j u + 2 s) - B_{j+2}
//   j = 0: z   = (c_0 u + s) - B_2, and the partials of <r, z>
// with u = C r (B_{j+2} = 0 when it does not exist).  st == nullptr runs the
// step unconditionally (restart); otherwise it is skipped once the
// iteration is off.
template <bool FINAL>
__global__ __launch_bounds__(kBlock) void k_poly_step(Crs<float> A, const float *__restrict__ b1,
                                                      const float *__restrict__ b2,
                                                      const float *__restrict__ r,
                                                      const float *__restrict__ C,
                                                      const PolyParams *poly, int j,
                                                      float *__restrict__ b0,
                                                      float *__restrict__ pcopy, Partials prz,
                                                      const MixedState *st) {
  __shared__ float prod[kNnzCap];
  __shared__ double lds[4];
  if (st && st->cur.done) return;
  const float cj = poly->c[j];
  const float tb = poly->two_over_b;
  double tpart = 0.0;
  for (int rb = blockIdx.x; rb < A.nrb; rb += gridDim.x) {
    tpart += crs_rowblock_epi<float, float>(A, b1, rb, prod, [&](int i, float t) -> double {
      const float ci = C[i];
      const float u = ci * r[i];
      const float sv = tb * (ci * t) - b1[i];
      const float bb = b2 ? b2[i] : 0.0f;
      if (FINAL) {
        const float z = (cj * u + sv) - bb;
        b0[i] = z;
        if (pcopy) pcopy[i] = z;
        return (double)r[i] * (double)z;
      }
      b0[i] = (cj * u + 2.0f * sv) - bb;
      return 0.0;
    });
  }
  if (FINAL) {
    const double bp = block_fold(tpart, lds);
    if (threadIdx.x == 0) prz.a[blockIdx.x] = bp;
  }
}

// beta = <r,z>_new / <r,z>; p32 = z + beta p32; stop test / reliable-update
// trigger on the recursive residual.  Writes `next`.
__global__ __launch_bounds__(kBlock) void k_mp_xpay_poly(int n, float *__restrict__ p,
                                                         const float *__restrict__ z, Partials prr,
                                                         Partials prz, int gz, MixedState *st,
                                                         History h) {
  __shared__ double lds[4];
  const MixedScalars c = st->cur;
  if (c.done) return;
  const double rr = fold_partials(prr.a, gridDim.x, lds);
  const double rz = fold_partials(prz.a, gz, lds);
  const double res = sqrt(rr) / st->nrm2_0;
  const double beta = rz / c.rho;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    const int it = c.iteration;
    MixedScalars s = c;
    s.iteration = it + 1;
    s.rho = rz;
    s.r2 = rr;
    s.residual = res;
    if (res <= st->tol || rr <= st->delta2 * c.r2_ref) s.done = 2;
    st->next = s;
    if (h.r2) h.r2[it + 1] = rr;
    if (h.rho) h.rho[it + 1] = rz;
    if (h.beta) h.beta[it] = beta;
  }
  const float b = (float)beta;
  Quads Q(n);
  for (; Q.q + (kQuadUnroll - 1) * Q.T < Q.full; Q.q += kQuadUnroll * Q.T) {
    float4 pv[kQuadUnroll], zv[kQuadUnroll];
#pragma unroll
    for (int u = 0; u < kQuadUnroll; u++) {
      pv[u] = v4(p)[Q.q + u * Q.T];
      zv[u] = v4(z)[Q.q + u * Q.T];
    }
#pragma unroll
    for (int u = 0; u < kQuadUnroll; u++) {
      float4 o;
      o.x = zv[u].x + b * pv[u].x;
      o.y = zv[u].y + b * pv[u].y;
      o.z = zv[u].z + b * pv[u].z;
      o.w = zv[u].w + b * pv[u].w;
      v4(p)[Q.q + u * Q.T] = o;
    }
  }
  for (; Q.q < Q.nquad; Q.q += Q.T)
    for (int64_t e = 4 * Q.q; e < 4 * Q.q + 4 && e < n; e++) p[e] = z[e] + b * p[e];
}

}  // namespace

hipError_t convert_f64_f32(int64_t n, const double *x, float *y, hipStream_t s) {
  int64_t g = (n + kBlock - 1) / kBlock;
  if (g > 8192) g = 8192;
  if (g < 1) g = 1;
  hipLaunchKernelGGL(k_convert, dim3((unsigned)g), dim3(kBlock), 0, s, n, x, y);
  return hipGetLastError();
}

hipError_t mp_accumulate(int n, double *x64, const float *x32, hipStream_t s) {
  hipLaunchKernelGGL(k_mp_accum, dim3(vec_grid(n)), dim3(kBlock), 0, s, n, x64, x32);
  return hipGetLastError();
}

hipError_t mp_residual(int n, const double *k, const double *q64, double *r64, Partials pa,
                       hipStream_t s) {
  hipLaunchKernelGGL(k_mp_residual, dim3(vec_grid(n)), dim3(kBlock), 0, s, n, k, q64, r64, pa);
  return hipGetLastError();
}

hipError_t mp_restart(int n, const double *r64, float *r32, float *x32, float *p32,
                      const float *C32, bool first, Partials pb, hipStream_t s) {
  if (C32)
    hipLaunchKernelGGL(k_mp_restart<true>, dim3(vec_grid4(n)), dim3(kBlock), 0, s, n, r64, r32, x32,
                       p32, C32, (int)first, pb);
  else
    hipLaunchKernelGGL(k_mp_restart<false>, dim3(vec_grid4(n)), dim3(kBlock), 0, s, n, r64, r32,
                       x32, p32, C32, (int)first, pb);
  return hipGetLastError();
}

hipError_t mp_restart_fold(int n, Partials prt, Partials prr, const double *rz, int gz, bool first,
                           MixedState *st, History h, hipStream_t s) {
  hipLaunchKernelGGL(k_mp_restart_fold, dim3(1), dim3(kBlock), 0, s, prt, vec_grid(n), prr,
                     vec_grid4(n), rz, gz, (int)first, st, h);
  return hipGetLastError();
}

hipError_t mp_restart_poly(int n, const double *r64, float *r32, float *x32, const float *C32,
                           const PolyParams *poly, float *bk, Partials prr, hipStream_t s) {
  hipLaunchKernelGGL(k_mp_restart_poly, dim3(vec_grid4(n)), dim3(kBlock), 0, s, n, r64, r32, x32,
                     C32, poly, bk, prr);
  return hipGetLastError();
}

hipError_t mp_poly_apply(const Crs<float> &A, const float *r32, const float *C32,
                         const PolyParams *poly, int degree, float *const w[3], float *pcopy,
                         Partials prz, const MixedState *st, hipStream_t s) {
  const dim3 grid(spmv_grid(A.nrb)), block(kBlock);
  for (int j = degree - 1; j >= 1; j--) {
    const float *b2 = (j + 2 <= degree) ? w[(j + 2) % 3] : nullptr;
    hipLaunchKernelGGL(k_poly_step<false>, grid, block, 0, s, A, w[(j + 1) % 3], b2, r32, C32, poly,
                       j, w[j % 3], (float *)nullptr, prz, st);
  }
  const float *b2 = degree >= 2 ? w[2] : nullptr;
  hipLaunchKernelGGL(k_poly_step<true>, grid, block, 0, s, A, w[1], b2, r32, C32, poly, 0, w[0],
                     pcopy, prz, st);
  return hipGetLastError();
}

hipError_t mp_update_poly(int n, int gpq, float *x32, float *r32, const float *p32,
                          const float *q32, const float *C32, const PolyParams *poly, float *bk,
                          Partials pq, Partials prr, MixedState *st, History h, hipStream_t s) {
  hipLaunchKernelGGL(k_mp_update_poly, dim3(vec_grid4(n)), dim3(kBlock), 0, s, n, gpq, x32, r32,
                     p32, q32, C32, poly, bk, pq, prr, st, h);
  return hipGetLastError();
}

hipError_t mp_xpay_poly(int n, float *p32, const float *z32, Partials prr, Partials prz, int gz,
                        MixedState *st, History h, hipStream_t s) {
  hipLaunchKernelGGL(k_mp_xpay_poly, dim3(vec_grid4(n)), dim3(kBlock), 0, s, n, p32, z32, prr, prz,
                     gz, st, h);
  return hipGetLastError();
}

hipError_t mp_matvec_dot_crs(const Crs<float> &A, const float *p, float *q, Partials pq,
                             MixedState *st, hipStream_t s) {
  hipLaunchKernelGGL(k_mp_matvec_dot_crs, dim3(spmv_grid(A.nrb)), dim3(kBlock), 0, s, A, p, q, pq,
                     st);
  return hipGetLastError();
}

hipError_t mp_update(int n, int gpq, float *x32, float *r32, const float *p32, const float *q32,
                     const float *C32, Partials pq, Partials pr, MixedState *st, History h,
                     hipStream_t s) {
  if (C32)
    hipLaunchKernelGGL(k_mp_update<true>, dim3(vec_grid4(n)), dim3(kBlock), 0, s, n, gpq, x32, r32,
                       p32, q32, C32, pq, pr, st, h);
  else
    hipLaunchKernelGGL(k_mp_update<false>, dim3(vec_grid4(n)), dim3(kBlock), 0, s, n, gpq, x32,
                       r32, p32, q32, C32, pq, pr, st, h);
  return hipGetLastError();
}

hipError_t mp_xpay(int n, float *p32, const float *r32, const float *C32, Partials pr,
                   MixedState *st, History h, hipStream_t s) {
  if (C32)
    hipLaunchKernelGGL(k_mp_xpay<true>, dim3(vec_grid4(n)), dim3(kBlock), 0, s, n, p32, r32, C32, pr,
                       st, h);
  else
    hipLaunchKernelGGL(k_mp_xpay<false>, dim3(vec_grid4(n)), dim3(kBlock), 0, s, n, p32, r32, C32,
                       pr, st, h);
  return hipGetLastError();
}

}  // namespace cghip
