/*
 * cghip_mlp.h -- C ABI of the FP32 network that maps matrix features to the
 * polynomial preconditioner's coefficients (no counterpart in the reference,
 * which has only the Jacobi preconditioner, Preconditioner.h:12-25).
 *
 * A fully connected network in -> h1 -> h2 -> out with ReLU after both
 * hidden layers, trained by full-batch Adam on the mean squared error
 * L = sum (Y - T)^2 / (rows * out).  Forward, backward and Adam are HIP
 * kernels; sums are in a fixed order, so results repeat bit for bit.
 *
 * Parameters are one float array, in this order:
 *   W1[in][h1] b1[h1] W2[h1][h2] b2[h2] W3[h2][out] b3[out]
 * with W[i][j] the weight from input i to output j.  Widths are 1..256.
 * Conventions as in cghip.h: int status, message via
 * cghip_mlp_last_error, one stream per handle, calls return after the
 * device work is complete, not thread-safe.
 */
#ifndef CGHIP_MLP_H
#define CGHIP_MLP_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct cghip_mlp cghip_mlp;

typedef struct {
  float lr;     /* step size alpha */
  float beta1;  /* first-moment decay */
  float beta2;  /* second-moment decay */
  float eps;
} cghip_adam_opts;

/* Number of parameters of a network of these widths (0 if invalid). */
int cghip_mlp_param_count(int in, int h1, int h2, int out);
/* max_rows bounds the rows of every later call. Parameters start at 0. */
int cghip_mlp_create(int device, int in, int h1, int h2, int out, int max_rows,
                     cghip_mlp **handle);
int cghip_mlp_destroy(cghip_mlp *m);
const char *cghip_mlp_last_error(const cghip_mlp *m);
/* Sets the parameters and resets the Adam moments and step count. */
int cghip_mlp_set_params(cghip_mlp *m, const float *params);
int cghip_mlp_get_params(cghip_mlp *m, float *params);
/* Y[rows][out] = network(X[rows][in]). */
int cghip_mlp_forward(cghip_mlp *m, const float *X, int rows, float *Y);
/* Gradient of the loss at the current parameters (grad has
 * cghip_mlp_param_count entries) and the loss itself. No update. */
int cghip_mlp_gradient(cghip_mlp *m, const float *X, const float *T, int rows,
                       float *grad, double *loss);
/* `steps` full-batch Adam steps on (X, T).  loss_history (nullable, steps
 * entries) receives the loss at the parameters each step starts from. */
int cghip_mlp_train(cghip_mlp *m, const float *X, const float *T, int rows,
                    int steps, const cghip_adam_opts *opts,
                    double *loss_history);

#ifdef __cplusplus
}
#endif
#endif
