// Test driver for the C++ host's learned-polynomial helpers (no GPU call):
//   learned_host ritz <m> <alpha...> <beta...>   -> "lo hi"
//   learned_host cheb <a> <b> <degree>            -> coefficients
//   learned_host model <path>                     -> "in h1 h2 out D fixed nparams rho", "degree c_k" per degree
//   learned_host decide <path> <y...>             -> "degree lower_end ms_per_iter cost_jacobi theta_min", costs
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "LearnedPoly.h"

int main(int argc, char **argv) {
  if (argc < 2) return 2;
  const std::string cmd = argv[1];
  if (cmd == "ritz") {
    const int m = std::atoi(argv[2]);
    std::vector<double> a(m), b(m);
    for (int i = 0; i < m; i++) a[i] = std::strtod(argv[3 + i], nullptr);
    for (int i = 0; i < m - 1; i++) b[i] = std::strtod(argv[3 + m + i], nullptr);
    double lo, hi;
    ritzExtremes(a.data(), b.data(), m, lo, hi);
    std::printf("%.17g %.17g\n", lo, hi);
  } else if (cmd == "cheb") {
    const auto c = chebResidualCoeffs(std::strtod(argv[2], nullptr), std::strtod(argv[3], nullptr),
                                      std::atoi(argv[4]));
    for (double v : c) std::printf("%.17g\n", v);
  } else if (cmd == "model") {
    PolyModel m;
    std::string err;
    if (!m.load(argv[2], err)) {
      std::printf("error: %s\n", err.c_str());
      return 1;
    }
    std::printf("%d %d %d %d %zu %d %zu %.17g\n", m.in, m.h1, m.h2, m.out, m.degrees.size(),
                m.fixedDegree, m.params.size(), m.rho_star);
    for (size_t j = 0; j < m.degrees.size(); j++)
      std::printf("%d %.17g\n", m.degrees[j], m.lowerEnd[j]);
  } else if (cmd == "decide") {
    PolyModel m;
    std::string err;
    if (!m.load(argv[2], err)) {
      std::printf("error: %s\n", err.c_str());
      return 1;
    }
    std::vector<float> y(m.out);
    for (int i = 0; i < m.out; i++) y[i] = std::strtof(argv[3 + i], nullptr);
    const LearnedDecision d = decide(m, y.data());
    std::printf("%d %.17g %.17g %.17g %.17g\n", d.degree, d.lowerEnd, d.msPerIter, d.costJacobi,
                d.thetaMin);
    for (double c : d.cost) std::printf("%.17g\n", c);
  }
  return 0;
}
