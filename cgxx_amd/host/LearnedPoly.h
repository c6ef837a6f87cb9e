// Mixed-precision solve with the learned polynomial preconditioner, driven
// from C++ through the C ABI (include/cghip.h, include/cghip_mlp.h).
//
// The same sequence as cgxx_amd/learned.py (solve_learned): device matrix
// statistics and log N (37 inputs, no solver steps), the network on the GPU
// for the smallest eigenvalue theta_min of D^-1 A and, per candidate degree,
// the iterations with Chebyshev on [c_k theta_min, b] and the time per
// iteration, and Jacobi's iterations and time per iteration; the cheapest
// predicted method runs, the polynomial within an iteration budget worth
// 1.25x Jacobi's predicted cost, with the Jacobi fallback judged by the FP64
// true residual.  The "fixed" mode is the rule without the network: 10 FP64
// Jacobi-PCG steps for the smallest Ritz value theta10, Chebyshev of one
// degree on [theta10 / rho*, b].  No reference counterpart.
#pragma once

#include <string>
#include <vector>

#include "cghip.h"
#include "cghip_mlp.h"

struct PolyModel {
  int in = 0, h1 = 0, h2 = 0, out = 0;
  std::vector<int> degrees;      // candidate degrees
  std::vector<double> lowerEnd;  // c_k: the degree's interval is [c_k theta_min, b]
  int fixedDegree = 8;           // the fixed rule: Chebyshev of this degree on
  double rho_star = 16.0;        // [theta10 / rho_star, b]
  std::vector<double> feat_mean, feat_std, feat_lo, feat_hi, out_mean, out_std;
  std::vector<float> params;
  // Reads the version-3 binary written by cgxx_amd.learned.export_bin;
  // false + err on a malformed file.
  bool load(const std::string &path, std::string &err);
};

// The choice from the network's standardised outputs (learned.decide):
// degree 0 = Jacobi; cost[i] the predicted ms of degrees[i], costJacobi
// Jacobi's; lowerEnd = c_k thetaMin for the chosen degree.
struct LearnedDecision {
  int degree = 0;
  double thetaMin = 0, lowerEnd = 0, msPerIter = 0, costJacobi = 0;
  std::vector<double> cost;
};
LearnedDecision decide(const PolyModel &model, const float *yStandardised);

struct LearnedResult {
  std::string method;      // poly, poly+jacobi, jacobi (refused), jacobi-selected
  int degree = 0;          // the polynomial's degree (0: none ran)
  int polyIterations = 0;  // inner iterations with the polynomial
  int jacobiIterations = 0;
  double trueResidual = 0;  // FP64 ||k - A x|| / ||k||
  double msMatvec = 0, msUpdate = 0, msXpay = 0;
};

class LearnedPoly {
 public:
  // mode "learned": the network's spectrum estimate and cost choice;
  // "fixed": Chebyshev on [theta_min(10 steps) / rho*, b], always the
  // polynomial.
  LearnedPoly(cghip_ctx *ctx, int device, const PolyModel &model, const std::string &mode);
  ~LearnedPoly();
  // Solves from X = 0 (the context holds the CRS matrix, the Jacobi
  // diagonal, K, and the FP32 data of cghip_prepare_fp32).  Returns false
  // with err set when a library call fails.
  bool solve(int n, double tol, int maxIter, LearnedResult &res, std::string &err);

 private:
  cghip_ctx *ctx_;
  int device_;
  PolyModel model_;
  std::string mode_;
  cghip_mlp *mlp_ = nullptr;
};

// Extreme eigenvalues of the Lanczos tridiagonal of m PCG steps
// (T_jj = 1/a_j + b_{j-1}/a_{j-1}, T_j,j+1 = sqrt(b_j)/a_j), by Sturm-sequence
// bisection.
void ritzExtremes(const double *alpha, const double *beta, int m, double &lo, double &hi);
// Chebyshev coefficients on [0, b] of the residual polynomial of degree + 1
// on [a, b] (the Chebyshev preconditioner).
std::vector<double> chebResidualCoeffs(double a, double b, int degree);
