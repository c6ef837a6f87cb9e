#include "LearnedPoly.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <fstream>

namespace {

constexpr int kRitzSteps = 10;
constexpr int kStats = 36;
constexpr double kPi = 3.14159265358979323846;

template <typename T>
bool readArray(std::ifstream &f, std::vector<T> &v, size_t n) {
  v.resize(n);
  f.read(reinterpret_cast<char *>(v.data()), (std::streamsize)(n * sizeof(T)));
  return (bool)f;
}

// Number of eigenvalues of the tridiagonal (d, e) below x (Sturm sequence).
int countBelow(const std::vector<double> &d, const std::vector<double> &e, double x) {
  int cnt = 0;
  double q = 1.0;
  for (size_t i = 0; i < d.size(); i++) {
    const double e2 = i ? e[i - 1] * e[i - 1] : 0.0;
    q = d[i] - x - (i ? e2 / q : 0.0);
    if (q == 0.0) q = -1e-300;
    if (q < 0) cnt++;
  }
  return cnt;
}

// k-th smallest eigenvalue (0-based) by bisection on the Gershgorin interval.
double eigenvalue(const std::vector<double> &d, const std::vector<double> &e, int k) {
  double lo = 1e300, hi = -1e300;
  for (size_t i = 0; i < d.size(); i++) {
    const double r = (i ? std::fabs(e[i - 1]) : 0.0) + (i + 1 < d.size() ? std::fabs(e[i]) : 0.0);
    lo = std::min(lo, d[i] - r);
    hi = std::max(hi, d[i] + r);
  }
  for (int it = 0; it < 200 && hi - lo > 1e-15 * std::max(std::fabs(lo), std::fabs(hi)); it++) {
    const double mid = 0.5 * (lo + hi);
    if (countBelow(d, e, mid) > k)
      hi = mid;
    else
      lo = mid;
  }
  return 0.5 * (lo + hi);
}

// Chebyshev coefficients (degree k, basis on [0, b]) of the polynomial with
// values v[j] at the points y_j = cos(pi (j + 0.5) / (k + 1)).
std::vector<double> dct(const std::vector<double> &v) {
  const int k1 = (int)v.size();
  std::vector<double> c(k1, 0.0);
  for (int m = 0; m < k1; m++) {
    double s = 0.0;
    for (int j = 0; j < k1; j++) s += v[j] * std::cos(m * kPi * (j + 0.5) / k1);
    c[m] = s * 2.0 / k1;
  }
  c[0] *= 0.5;
  return c;
}

// T_n(s) by the three-term recurrence (any real s).
double chebT(int n, double s) {
  double t0 = 1.0, t1 = s;
  if (n == 0) return t0;
  for (int i = 1; i < n; i++) {
    const double t2 = 2.0 * s * t1 - t0;
    t0 = t1;
    t1 = t2;
  }
  return t1;
}

}  // namespace

bool PolyModel::load(const std::string &path, std::string &err) {
  std::ifstream f(path, std::ios::binary);
  if (!f) {
    err = "cannot open model file " + path;
    return false;
  }
  char magic[8];
  f.read(magic, 8);
  int32_t hdr[6];
  f.read(reinterpret_cast<char *>(hdr), sizeof(hdr));
  if (!f || std::memcmp(magic, "CGHIPMLP", 8) != 0 || hdr[0] != 3) {
    err = "not a version-3 CGHIPMLP model file: " + path;
    return false;
  }
  in = hdr[1], h1 = hdr[2], h2 = hdr[3], out = hdr[4];
  const int D = hdr[5];
  if (D < 1 || D > 16) {
    err = "truncated or inconsistent model file " + path;
    return false;
  }
  std::vector<int32_t> deg(D + 1);
  f.read(reinterpret_cast<char *>(deg.data()), sizeof(int32_t) * (D + 1));
  degrees.assign(deg.begin(), deg.begin() + D);
  fixedDegree = deg[D];
  bool shape = f && in == kStats + 1 && out == 2 * D + 3 && fixedDegree >= 1 && fixedDegree <= 16;
  for (int k : degrees) shape = shape && k >= 1 && k <= 16;
  if (!shape) {
    err = "model shape does not match the 37 features and 2 D + 3 outputs";
    return false;
  }
  f.read(reinterpret_cast<char *>(&rho_star), sizeof(double));
  bool ok = readArray(f, lowerEnd, D) && readArray(f, feat_mean, in) &&
            readArray(f, feat_std, in) && readArray(f, feat_lo, in) && readArray(f, feat_hi, in) &&
            readArray(f, out_mean, out) && readArray(f, out_std, out);
  for (double c : lowerEnd) ok = ok && c > 0;
  int32_t np = 0;
  f.read(reinterpret_cast<char *>(&np), sizeof(np));
  ok = ok && f && np == cghip_mlp_param_count(in, h1, h2, out) && readArray(f, params, np);
  if (!ok) err = "truncated or inconsistent model file " + path;
  return ok;
}

LearnedDecision decide(const PolyModel &m, const float *yn) {
  const int D = (int)m.degrees.size();
  auto y = [&](int i) { return (double)yn[i] * m.out_std[i] + m.out_mean[i]; };
  LearnedDecision d;
  d.thetaMin = std::exp(y(0));
  d.costJacobi = std::exp(y(1 + 2 * D) + y(2 + 2 * D));
  double best = d.costJacobi;
  d.cost.resize(D);
  for (int j = 0; j < D; j++) {
    d.cost[j] = std::exp(y(1 + j) + y(1 + D + j));
    if (d.cost[j] < best) {
      best = d.cost[j];
      d.degree = m.degrees[j];
      d.lowerEnd = m.lowerEnd[j] * d.thetaMin;
      d.msPerIter = std::exp(y(1 + D + j));
    }
  }
  return d;
}

void ritzExtremes(const double *alpha, const double *beta, int m, double &lo, double &hi) {
  std::vector<double> d(m), e(m > 1 ? m - 1 : 0);
  for (int j = 0; j < m; j++) {
    d[j] = 1.0 / alpha[j] + (j ? beta[j - 1] / alpha[j - 1] : 0.0);
    if (j + 1 < m) e[j] = std::sqrt(beta[j]) / alpha[j];
  }
  lo = eigenvalue(d, e, 0);
  hi = eigenvalue(d, e, m - 1);
}

std::vector<double> chebResidualCoeffs(double a, double b, int degree) {
  const int k1 = degree + 1;
  const double s0 = (b + a) / (b - a);
  const double t0 = chebT(k1, s0);
  std::vector<double> v(k1);
  for (int j = 0; j < k1; j++) {
    const double y = std::cos(kPi * (j + 0.5) / k1);
    const double lam = (y + 1.0) * 0.5 * b;
    const double s = (b + a - 2.0 * lam) / (b - a);
    v[j] = (1.0 - chebT(k1, s) / t0) / lam;
  }
  return dct(v);
}

LearnedPoly::LearnedPoly(cghip_ctx *ctx, int device, const PolyModel &model,
                         const std::string &mode)
    : ctx_(ctx), device_(device), model_(model), mode_(mode) {}

LearnedPoly::~LearnedPoly() {
  if (mlp_) cghip_mlp_destroy(mlp_);
}

bool LearnedPoly::solve(int n, double tol, int maxIter, LearnedResult &res, std::string &err) {
  auto fail = [&](const char *what) {
    err = std::string(what) + ": " + cghip_last_error(ctx_);
    return false;
  };
  // ---- the network's inputs: matrix statistics and log N ----
  std::vector<double> f(model_.in, 0.0);
  if (cghip_matrix_stats(ctx_, f.data(), kStats)) return fail("matrix statistics");
  const double b = 1.0 + f[7];  // 1 + max Gershgorin radius of D^-1 A
  f[kStats] = std::log((double)n);

  // ---- choice and coefficients ----
  int k = model_.fixedDegree;
  std::vector<double> coeffs;
  bool usePoly = true;
  int budgetIters = std::max(1, maxIter / (k + 1));
  if (mode_ == "fixed") {
    // the rule without the network estimates the spectrum: 10 FP64
    // Jacobi-PCG steps for the smallest Ritz value
    if (cghip_fill_vector(ctx_, CGHIP_VEC_X, 0.0)) return fail("fill");
    cghip_solve_opts so{CGHIP_PRECOND_JACOBI, kRitzSteps, 1e-300, 0, CGHIP_PROFILE_NONE};
    cghip_solve_info si{};
    if (cghip_solve(ctx_, &so, &si)) return fail("10 PCG steps");
    double alpha[kRitzSteps], beta[kRitzSteps];
    int w = 0;
    if (si.iterations < kRitzSteps || cghip_history(ctx_, 3, alpha, kRitzSteps, &w) ||
        cghip_history(ctx_, 4, beta, kRitzSteps, &w))
      return fail("PCG history");
    double lo10, hi10;
    ritzExtremes(alpha, beta, kRitzSteps, lo10, hi10);
    coeffs = chebResidualCoeffs(std::min(lo10 / model_.rho_star, 0.25 * b), b, k);
  } else {
    if (!mlp_ && cghip_mlp_create(device_, model_.in, model_.h1, model_.h2, model_.out, 1, &mlp_)) {
      err = std::string("network: ") + cghip_mlp_last_error(nullptr);
      return false;
    }
    if (cghip_mlp_set_params(mlp_, model_.params.data())) {
      err = std::string("network parameters: ") + cghip_mlp_last_error(mlp_);
      return false;
    }
    std::vector<float> xin(model_.in), yout(model_.out);
    for (int i = 0; i < model_.in; i++) {
      const double c = std::min(std::max(f[i], model_.feat_lo[i]), model_.feat_hi[i]);
      xin[i] = (float)((c - model_.feat_mean[i]) / model_.feat_std[i]);
    }
    if (cghip_mlp_forward(mlp_, xin.data(), 1, yout.data())) {
      err = std::string("network forward: ") + cghip_mlp_last_error(mlp_);
      return false;
    }
    const LearnedDecision d = decide(model_, yout.data());
    usePoly = d.degree > 0;
    if (usePoly) {
      k = d.degree;
      coeffs = chebResidualCoeffs(std::min(d.lowerEnd, 0.25 * b), b, k);
      budgetIters = std::max(1, (int)std::ceil(1.25 * d.costJacobi / d.msPerIter));
    }
  }
  const bool accepted = usePoly && cghip_set_poly(ctx_, k, coeffs.data(), b) == CGHIP_OK;

  // ---- solve, with the Jacobi fallback ----
  if (cghip_fill_vector(ctx_, CGHIP_VEC_X, 0.0)) return fail("fill");
  double rel = 1.0;
  res = LearnedResult{};
  res.method = !usePoly ? "jacobi-selected" : accepted ? "poly" : "jacobi (refused)";
  res.degree = accepted ? k : 0;
  cghip_mixed_info mi{};
  if (accepted) {
    cghip_mixed_opts mo{CGHIP_PRECOND_POLY, std::min(budgetIters, maxIter), tol, 0.0, 0,
                        CGHIP_PROFILE_NONE};
    if (cghip_solve_mixed(ctx_, &mo, &mi)) return fail("polynomial solve");
    res.polyIterations = mi.iterations;
    rel = mi.true_residual;
    if (!(rel <= tol)) res.method = "poly+jacobi";
  }
  // !(rel <= tol) also catches a NaN residual; a non-finite one leaves no
  // usable X, so Jacobi then starts again from X = 0.
  if (!(rel <= tol)) {
    if (!std::isfinite(rel)) {
      if (cghip_fill_vector(ctx_, CGHIP_VEC_X, 0.0)) return fail("fill");
      rel = 1.0;
    }
    cghip_mixed_opts mo{CGHIP_PRECOND_JACOBI, maxIter, tol / rel, 0.0, 0, CGHIP_PROFILE_NONE};
    mi = cghip_mixed_info{};
    if (cghip_solve_mixed(ctx_, &mo, &mi)) return fail("Jacobi solve");
    res.jacobiIterations = mi.iterations;
    rel *= mi.true_residual;
  }
  res.trueResidual = rel;
  return true;
}
