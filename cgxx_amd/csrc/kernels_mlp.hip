// Small fully connected network in FP32 for gfx950: in -> h1 -> h2 -> out,
// ReLU after the two hidden layers, mean-squared-error loss, full-batch
// gradients and Adam (Kingma & Ba, "Adam: A Method for Stochastic
// Optimization", ICLR 2015, Algorithm 1).
//
// The network is tiny (a few ten thousand parameters) and the batches are a
// few thousand rows, so every launch is latency-bound: the kernels keep one
// sample per block (forward) or a strided set of samples per block
// (backward) with all activations in LDS and the weights read through the
// caches.  No matrix cores: the products are 40..128 wide.
//
// Determinism: every sum has a fixed order.  A forward dot adds inputs in
// ascending index order; the backward kernel accumulates its block's
// samples in ascending order into the block's own slice of a partial-
// gradient buffer (each entry owned by one thread, no atomics); the Adam
// kernel adds the G slices in ascending block order.
//
// Parameter layout (one contiguous float array):
//   W1[in][h1] b1[h1] W2[h1][h2] b2[h2] W3[h2][out] b3[out]
// (W[i][j] = weight from input i to output j, row-major).

#include "mlp.h"

#include <hip/hip_runtime.h>

#pragma clang fp contract(off)

namespace cghip {
namespace {

constexpr int kThreads = 256;

struct Offsets {
  int w1, b1, w2, b2, w3, b3, total;
};

__host__ __device__ inline Offsets offsets(const MlpDims &d) {
  Offsets o;
  o.w1 = 0;
  o.b1 = o.w1 + d.in * d.h1;
  o.w2 = o.b1 + d.h1;
  o.b2 = o.w2 + d.h1 * d.h2;
  o.w3 = o.b2 + d.h2;
  o.b3 = o.w3 + d.h2 * d.out;
  o.total = o.b3 + d.out;
  return o;
}

// y[j] = b[j] + sum_i x[i] W[i][j] for j < nout (thread j), optional ReLU.
__device__ inline void dense_layer(const float *__restrict__ W, const float *__restrict__ b,
                                   const float *x, int nin, int nout, bool relu, float *y) {
  const int j = threadIdx.x;
  if (j < nout) {
    float acc = b[j];
    for (int i = 0; i < nin; i++) acc = acc + x[i] * W[(size_t)i * nout + j];
    y[j] = relu ? fmaxf(acc, 0.0f) : acc;
  }
}

// One block per sample: activations a1, a2 and the output y.
__global__ __launch_bounds__(kThreads) void k_mlp_forward(MlpDims d, const float *__restrict__ P,
                                                          const float *__restrict__ X, int rows,
                                                          float *__restrict__ A1,
                                                          float *__restrict__ A2,
                                                          float *__restrict__ Y) {
  __shared__ float x[kMlpMaxWidth], a1[kMlpMaxWidth], a2[kMlpMaxWidth], y[kMlpMaxWidth];
  const int s = blockIdx.x;
  if (s >= rows) return;
  const Offsets o = offsets(d);
  const int t = threadIdx.x;
  if (t < d.in) x[t] = X[(size_t)s * d.in + t];
  __syncthreads();
  dense_layer(P + o.w1, P + o.b1, x, d.in, d.h1, true, a1);
  __syncthreads();
  dense_layer(P + o.w2, P + o.b2, a1, d.h1, d.h2, true, a2);
  __syncthreads();
  dense_layer(P + o.w3, P + o.b3, a2, d.h2, d.out, false, y);
  if (A1 && t < d.h1) A1[(size_t)s * d.h1 + t] = a1[t];
  if (A2 && t < d.h2) A2[(size_t)s * d.h2 + t] = a2[t];
  if (t < d.out) Y[(size_t)s * d.out + t] = y[t];
}

// Loss L = sum_{s,o} (Y - T)^2 / (rows * out).  Block g accumulates the
// gradients of samples g, g + G, ... (ascending) into Gp[g * total + k] and
// its loss share into Lp[g].
__global__ __launch_bounds__(kThreads) void k_mlp_backward(MlpDims d, const float *__restrict__ P,
                                                           const float *__restrict__ X,
                                                           const float *__restrict__ A1,
                                                           const float *__restrict__ A2,
                                                           const float *__restrict__ Y,
                                                           const float *__restrict__ T, int rows,
                                                           float *__restrict__ Gp,
                                                           double *__restrict__ Lp) {
  __shared__ float x[kMlpMaxWidth], a1[kMlpMaxWidth], a2[kMlpMaxWidth];
  __shared__ float dy[kMlpMaxWidth], dz2[kMlpMaxWidth], dz1[kMlpMaxWidth];
  __shared__ double lsum[kThreads / 64];
  const Offsets o = offsets(d);
  const int t = threadIdx.x;
  const int G = gridDim.x;
  float *gp = Gp + (size_t)blockIdx.x * o.total;
  const float scale = 2.0f / ((float)rows * (float)d.out);
  double loss = 0.0;
  bool first = true;
  for (int s = blockIdx.x; s < rows; s += G) {
    if (t < d.in) x[t] = X[(size_t)s * d.in + t];
    if (t < d.h1) a1[t] = A1[(size_t)s * d.h1 + t];
    if (t < d.h2) a2[t] = A2[(size_t)s * d.h2 + t];
    if (t < d.out) {
      const float e = Y[(size_t)s * d.out + t] - T[(size_t)s * d.out + t];
      dy[t] = scale * e;
      loss += (double)e * (double)e;
    }
    __syncthreads();
    // dz2[j] = (sum_o W3[j][o] dy[o]) * [a2[j] > 0]
    if (t < d.h2) {
      float acc = 0.0f;
      const float *w = P + o.w3 + (size_t)t * d.out;
      for (int k = 0; k < d.out; k++) acc = acc + w[k] * dy[k];
      dz2[t] = a2[t] > 0.0f ? acc : 0.0f;
    }
    __syncthreads();
    // dz1[i] = (sum_j W2[i][j] dz2[j]) * [a1[i] > 0]
    if (t < d.h1) {
      float acc = 0.0f;
      const float *w = P + o.w2 + (size_t)t * d.h2;
      for (int k = 0; k < d.h2; k++) acc = acc + w[k] * dz2[k];
      dz1[t] = a1[t] > 0.0f ? acc : 0.0f;
    }
    __syncthreads();
    // Outer products into the block's partial gradient.
    auto put = [&](int k, float v) { gp[k] = first ? v : gp[k] + v; };
    for (int k = t; k < d.in * d.h1; k += kThreads) put(o.w1 + k, x[k / d.h1] * dz1[k % d.h1]);
    for (int k = t; k < d.h1; k += kThreads) put(o.b1 + k, dz1[k]);
    for (int k = t; k < d.h1 * d.h2; k += kThreads) put(o.w2 + k, a1[k / d.h2] * dz2[k % d.h2]);
    for (int k = t; k < d.h2; k += kThreads) put(o.b2 + k, dz2[k]);
    for (int k = t; k < d.h2 * d.out; k += kThreads) put(o.w3 + k, a2[k / d.out] * dy[k % d.out]);
    for (int k = t; k < d.out; k += kThreads) put(o.b3 + k, dy[k]);
    first = false;
    __syncthreads();
  }
  if (first) {  // a block without samples (rows < G) contributes zeros
    for (int k = t; k < o.total; k += kThreads) gp[k] = 0.0f;
  }
  // loss: lanes of wave 0 hold the per-output sums; fold in a fixed order
  for (int off = 32; off >= 1; off >>= 1) loss += __shfl_down(loss, off, 64);
  if ((t & 63) == 0) lsum[t >> 6] = loss;
  __syncthreads();
  if (t == 0) Lp[blockIdx.x] = ((lsum[0] + lsum[1]) + lsum[2]) + lsum[3];
}

// g = sum_{b < G} Gp[b][k] (ascending b); optional copy of g; Adam step.
__global__ __launch_bounds__(kThreads) void k_mlp_adam(int total, int G, const float *__restrict__ Gp,
                                                       float *__restrict__ grad, float *__restrict__ P,
                                                       float *__restrict__ M, float *__restrict__ V,
                                                       AdamStep a) {
  const int k = blockIdx.x * kThreads + threadIdx.x;
  if (k >= total) return;
  float g = 0.0f;
  for (int b = 0; b < G; b++) g = g + Gp[(size_t)b * total + k];
  if (grad) grad[k] = g;
  if (!a.apply) return;
  const float m = a.beta1 * M[k] + (1.0f - a.beta1) * g;
  const float v = a.beta2 * V[k] + (1.0f - a.beta2) * (g * g);
  M[k] = m;
  V[k] = v;
  const float mhat = m * a.corr1;  // corr1 = 1 / (1 - beta1^t)
  const float vhat = v * a.corr2;  // corr2 = 1 / (1 - beta2^t)
  P[k] = P[k] - a.lr * mhat / (sqrtf(vhat) + a.eps);
}

// One block: loss = sum_b Lp[b] (ascending b) / (rows * out).
__global__ void k_mlp_loss(int G, const double *__restrict__ Lp, double denom, double *out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    double s = 0.0;
    for (int b = 0; b < G; b++) s += Lp[b];
    *out = s / denom;
  }
}

}  // namespace

int mlp_param_count(const MlpDims &d) { return offsets(d).total; }

hipError_t mlp_forward(const MlpDims &d, const float *P, const float *X, int rows, float *A1,
                       float *A2, float *Y, hipStream_t s) {
  if (rows <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_mlp_forward, dim3(rows), dim3(kThreads), 0, s, d, P, X, rows, A1, A2, Y);
  return hipGetLastError();
}

hipError_t mlp_backward(const MlpDims &d, const float *P, const float *X, const float *A1,
                        const float *A2, const float *Y, const float *T, int rows, int G, float *Gp,
                        double *Lp, hipStream_t s) {
  hipLaunchKernelGGL(k_mlp_backward, dim3(G), dim3(kThreads), 0, s, d, P, X, A1, A2, Y, T, rows, Gp,
                     Lp);
  return hipGetLastError();
}

hipError_t mlp_adam(int total, int G, const float *Gp, float *grad, float *P, float *M, float *V,
                    const AdamStep &a, hipStream_t s) {
  hipLaunchKernelGGL(k_mlp_adam, dim3((total + kThreads - 1) / kThreads), dim3(kThreads), 0, s,
                     total, G, Gp, grad, P, M, V, a);
  return hipGetLastError();
}

hipError_t mlp_loss(int G, const double *Lp, double denom, double *out, hipStream_t s) {
  hipLaunchKernelGGL(k_mlp_loss, dim3(1), dim3(64), 0, s, G, Lp, denom, out);
  return hipGetLastError();
}

}  // namespace cghip
