// Internal interface of the FP32 network kernels (kernels_mlp.hip).
#pragma once

#include <hip/hip_runtime.h>

namespace cghip {

constexpr int kMlpMaxWidth = 256;  // every layer width <= one block of threads

struct MlpDims {
  int in, h1, h2, out;
};

struct AdamStep {
  float lr, beta1, beta2, eps;
  float corr1, corr2;  // 1 / (1 - beta1^t), 1 / (1 - beta2^t)
  int apply;           // 0: only fold the gradient
};

int mlp_param_count(const MlpDims &d);

// A1/A2 (rows x h1 / h2, nullable) keep the activations for the backward pass.
hipError_t mlp_forward(const MlpDims &d, const float *P, const float *X, int rows, float *A1,
                       float *A2, float *Y, hipStream_t s);
// Partial gradients Gp[G][total] and loss shares Lp[G] of the MSE loss.
hipError_t mlp_backward(const MlpDims &d, const float *P, const float *X, const float *A1,
                        const float *A2, const float *Y, const float *T, int rows, int G, float *Gp,
                        double *Lp, hipStream_t s);
// Folds Gp into the gradient (copied to `grad` when non-null) and, when
// a.apply, takes one Adam step on P with moments M, V.
hipError_t mlp_adam(int total, int G, const float *Gp, float *grad, float *P, float *M, float *V,
                    const AdamStep &a, hipStream_t s);
hipError_t mlp_loss(int G, const double *Lp, double denom, double *out, hipStream_t s);

}  // namespace cghip
