// CGHIP: the MI355X backend of the CG driver.
//
// Replaces the reference's CUDA backend (cuda/CGCUDA.cu, cuda/CGCUDABase.h)
// with calls into the C ABI of include/cghip.h; every kernel is a HIP
// kernel for gfx950 in libcghip.so.  Each hook maps to one ABI call and,
// like the reference's hooks, returns after the device work is complete.
// A failing call prints "HIP error: <message>" and exits with status 1
// (cuda/utils.h:10-15 does the same with "CUDA error").
//
// By default solve() runs the whole PCG loop on the device (solveFused):
// three kernels per iteration, the scalars (alpha, beta, the stop test)
// computed on the device, kernel times measured with HIP events and
// charged to the reference's categories
// (SpMV + <p,q> -> MatVec; r update + <r,r> + <r,z> -> axpy; x update,
// Jacobi and p update -> xpay; vectorDot and Preconditioner are fused into
// those and report 0).  Environment (in addition to the reference's CG_* variables):
//   CG_HIP_DEVICE   device index (default 0)
//   CG_HIP_FUSED    0: run the reference loop hook by hook instead
//   CG_HIP_DOT      tree (default) | serial: summation order of the dot hook;
//                   serial reproduces SerialCG's left-to-right order
//   CG_HIP_GRAPH    1: replay the fused iterations from a HIP graph (fastest;
//                   the per-category kernel times are then not measured and
//                   print as 0).  Default: one iteration at a time with every
//                   kernel timed by HIP events (one host round trip per
//                   iteration).
//   CG_HIP_MIXED    1: FP32-inner CG with FP64 reliable updates (CRS or COO;
//                   with Jacobi on the scaled operator D^-1/2 A D^-1/2); the
//                   reported residual is the FP64 true relative residual.
//   CG_HIP_POLY     none (default) | learned | fixed: with CG_HIP_MIXED=1 and
//                   Jacobi, the polynomial preconditioner chosen by the
//                   trained network from the matrix statistics (degree and
//                   interval from its predicted theta_min, or Jacobi) or by
//                   the fixed Chebyshev rule (10 FP64 PCG steps for theta10),
//                   both with the Jacobi fallback.
//   CG_HIP_MODEL    model file for CG_HIP_POLY (default: ../data/poly_mlp.bin
//                   next to this executable).
#include <cerrno>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <iomanip>
#include <iostream>
#include <string>
#include <vector>

#include <limits.h>
#include <unistd.h>

#include <memory>

#include "CG.h"
#include "LearnedPoly.h"
#include "cghip.h"

namespace {

[[noreturn]] void hip_fail(cghip_ctx *ctx, const char *what) {
  std::cerr << "HIP error: " << what << ": " << cghip_last_error(ctx) << std::endl;
  std::exit(1);
}

int env_int(const char *name, int def) {
  const char *v = std::getenv(name);
  if (!v || !*v) return def;
  char *end = nullptr;
  errno = 0;
  const long r = std::strtol(v, &end, 0);
  if (errno || *end || r < 0 || r > 0x7fffffffL) {
    std::cerr << "Invalid value for " << name << "!" << std::endl;
    std::exit(1);
  }
  return (int)r;
}

}  // namespace

class CGHIP : public CG {
  cghip_ctx *ctx = nullptr;
  int device = 0;
  bool fused = true;
  bool graph = false;
  int dotMode = CGHIP_DOT_TREE;
  bool mixed = false;
  std::string polyMode = "none";
  std::string modelPath;
  std::unique_ptr<LearnedPoly> learned;

  // What the last solve ran, for the solve record.
  std::string lastLoop = "hooks";
  cghip_mixed_info lastMixed{};
  LearnedResult lastLearned;

  bool solveMixed(int maxIterations, floatType tolerance, FusedResult &result);
  void recordBackend(std::ostream &os) override;
  // COO runs on the device CRS; MI355X HBM3E peak 8 TB/s.
  double matvecBytes() const override {
    const double n = N, z = nz;
    return getMatrixFormat() == MatrixFormatELL ? 12.0 * z + 20.0 * n
                                                : 12.0 * z + 4.0 * (n + 1) + 16.0 * n;
  }
  double peakGBs() const override { return 8000.0; }
  // ||r_i|| / ||r_0|| from the device's r2 history of the last solve.
  void residualHistory(int iterations, std::vector<double> &out) {
    std::vector<double> r2((size_t)iterations + 1);
    int got = 0;
    chk(cghip_history(ctx, 0, r2.data(), iterations + 1, &got), "history");
    out.resize((size_t)got);
    for (int i = 0; i < got; i++) out[i] = std::sqrt(r2[i] / r2[0]);
  }

  void chk(int rc, const char *what) {
    if (rc != CGHIP_OK) hip_fail(ctx, what);
  }

  // COO matrices run on the CRS kernels (see doTransferTo).
  bool supportsMatrixFormat(MatrixFormat format) override {
    return format == MatrixFormatCOO || format == MatrixFormatCRS || format == MatrixFormatELL;
  }
  bool supportsPreconditioner(Preconditioner p) override { return p == PreconditionerJacobi; }
  bool needsTransfer() override { return true; }

  void doTransferTo() override;
  void doTransferFrom() override;

  void cpy(Vector dst, Vector src) override { chk(cghip_cpy(ctx, dst, src), "cpy"); }
  void matvecKernel(Vector x_, Vector y_) override { chk(cghip_matvec(ctx, x_, y_), "matvec"); }
  void axpyKernel(floatType a, Vector x_, Vector y_) override {
    chk(cghip_axpy(ctx, a, x_, y_), "axpy");
  }
  void xpayKernel(Vector x_, floatType a, Vector y_) override {
    chk(cghip_xpay(ctx, x_, a, y_), "xpay");
  }
  floatType vectorDotKernel(Vector a, Vector b) override {
    double r = 0;
    chk(cghip_vector_dot(ctx, a, b, &r), "vectorDot");
    return r;
  }
  void applyPreconditionerKernel(Vector x_, Vector y_) override {
    chk(cghip_apply_preconditioner(ctx, x_, y_), "applyPreconditioner");
  }
  bool solveFused(int maxIterations, floatType tolerance, FusedResult &result) override;

public:
  CGHIP() : CG(MatrixFormatELL, PreconditionerJacobi) {}
  ~CGHIP() override {
    if (ctx) cghip_destroy(ctx);
  }

  void parseEnvironment() override {
    CG::parseEnvironment();
    device = env_int("CG_HIP_DEVICE", 0);
    fused = env_int("CG_HIP_FUSED", 1) != 0;
    graph = env_int("CG_HIP_GRAPH", 0) != 0;
    mixed = env_int("CG_HIP_MIXED", 0) != 0;
    if (const char *v = std::getenv("CG_HIP_POLY")) {
      const std::string s(v);
      if (s == "none" || s == "learned" || s == "fixed")
        polyMode = s;
      else if (!s.empty()) {
        std::cerr << "Invalid value for CG_HIP_POLY! (none, learned, or fixed)" << std::endl;
        std::exit(1);
      }
    }
    if (const char *v = std::getenv("CG_HIP_MODEL")) modelPath = v;
    if (polyMode != "none" && !mixed) {
      std::cerr << "CG_HIP_POLY needs CG_HIP_MIXED=1!" << std::endl;
      std::exit(1);
    }
    if (const char *v = std::getenv("CG_HIP_DOT")) {
      const std::string s(v);
      if (s == "tree")
        dotMode = CGHIP_DOT_TREE;
      else if (s == "serial")
        dotMode = CGHIP_DOT_SERIAL;
      else if (!s.empty()) {
        std::cerr << "Invalid value for CG_HIP_DOT! (tree, or serial)" << std::endl;
        std::exit(1);
      }
    }
  }

  void init(const char *matrixFile) override {
    // Like CGCUDA::init (cuda/CGCUDA.cu:29-36): pick the device first.
    if (cghip_create(device, &ctx) != CGHIP_OK) hip_fail(nullptr, "create");
    chk(cghip_set_dot_mode(ctx, dotMode), "set_dot_mode");
    CG::init(matrixFile);
  }
};

void CGHIP::doTransferTo() {
  chk(cghip_set_size(ctx, N), "set_size");
  if (matrixFormat == MatrixFormatCOO) {
    // SerialCG's COO SpMV (SerialCG.cpp:79-85) adds each row's products from
    // y = 0 in COO order, which is the CRS kernels' left-to-right order over
    // the stable CRS of the same entries.  Rows in ascending order (general
    // files written by rows) make J and V that CRS as they are, with row
    // pointers from the row counts; other orders (symmetric files) go
    // through the stable parallel conversion first.
    const MatrixCOO &coo = *matrixCOO;
    if (coo.rowsSorted) {
      std::vector<int> ptr((size_t)N + 1, 0);
      for (int i = 0; i < N; i++) ptr[i + 1] = ptr[i] + coo.nzPerRow[i];
      chk(cghip_upload_crs(ctx, N, nz, ptr.data(), coo.J.get(), coo.V.get()), "upload_crs");
    } else {
      MatrixCRS crs;
      crs.convert(coo);
      chk(cghip_upload_crs(ctx, N, nz, crs.ptr, crs.index, crs.value), "upload_crs");
      crs.deallocate();
    }
    chk(cghip_set_format(ctx, CGHIP_FORMAT_CRS), "set_format");
  } else if (matrixFormat == MatrixFormatCRS) {
    chk(cghip_upload_crs(ctx, N, nz, matrixCRS->ptr, matrixCRS->index, matrixCRS->value),
        "upload_crs");
    chk(cghip_set_format(ctx, CGHIP_FORMAT_CRS), "set_format");
  } else {
    chk(cghip_upload_ell(ctx, N, matrixELL->maxNz, matrixELL->length, matrixELL->index,
                         matrixELL->data),
        "upload_ell");
    chk(cghip_set_format(ctx, CGHIP_FORMAT_ELL), "set_format");
  }
  if (preconditioner == PreconditionerJacobi) chk(cghip_upload_jacobi(ctx, jacobi->C), "upload_jacobi");
  chk(cghip_upload_vector(ctx, CGHIP_VEC_K, k), "upload k");
  chk(cghip_upload_vector(ctx, CGHIP_VEC_X, x), "upload x");
  if (mixed) {
    if (matrixFormat == MatrixFormatELL) {
      std::cerr << "CG_HIP_MIXED needs CG_MATRIX_FORMAT=CRS (or COO)!" << std::endl;
      std::exit(1);
    }
    chk(cghip_prepare_fp32(ctx), "prepare_fp32");
    if (polyMode != "none" && !learned) {
      if (preconditioner != PreconditionerJacobi) {
        std::cerr << "CG_HIP_POLY needs CG_PRECONDITIONER=jacobi!" << std::endl;
        std::exit(1);
      }
      std::string path = modelPath;
      if (path.empty()) {
        char exe[PATH_MAX];
        const ssize_t len = readlink("/proc/self/exe", exe, sizeof(exe) - 1);
        std::string dir = len > 0 ? std::string(exe, (size_t)len) : std::string(".");
        dir = dir.substr(0, dir.find_last_of('/'));
        path = dir + "/../data/poly_mlp.bin";
      }
      PolyModel model;
      std::string err;
      if (!model.load(path, err)) {
        std::cerr << "HIP error: " << err << std::endl;
        std::exit(1);
      }
      learned.reset(new LearnedPoly(ctx, device, model, polyMode));
    }
  }
}

bool CGHIP::solveMixed(int maxIterations, floatType tolerance, FusedResult &result) {
  if (learned) {
    LearnedResult r;
    std::string err;
    if (!learned->solve(N, tolerance, maxIterations, r, err)) {
      std::cerr << "HIP error: " << err << std::endl;
      std::exit(1);
    }
    std::cout << "HIP mixed precision, polynomial (" << polyMode << "): " << r.method
              << (r.degree ? ", degree " + std::to_string(r.degree) : std::string()) << ", "
              << r.polyIterations << " polynomial + " << r.jacobiIterations
              << " Jacobi inner iterations" << std::endl;
    lastLoop = "mixed";
    lastLearned = r;
    result.iterations = r.polyIterations + r.jacobiIterations;
    result.residual = r.trueResidual;
    return true;
  }
  cghip_mixed_opts o{};
  o.precond = preconditioner == PreconditionerJacobi ? CGHIP_PRECOND_JACOBI : CGHIP_PRECOND_NONE;
  o.max_iter = maxIterations;
  o.tol = tolerance;
  o.delta = 0.0;
  o.batch = 0;
  o.profile = graph ? CGHIP_PROFILE_NONE : CGHIP_PROFILE_KERNELS;
  cghip_mixed_info info{};
  chk(cghip_solve_mixed(ctx, &o, &info), "solve_mixed");
  std::cout << "HIP mixed precision: " << info.updates << " reliable updates" << std::endl;
  lastLoop = "mixed";
  lastMixed = info;
  result.iterations = info.iterations;
  result.residual = info.true_residual;
  residualHistory(info.iterations, result.residuals);  // the inner recursion
  result.matvec = (info.ms_matvec + info.ms_refresh) * 1e-3;
  result.axpy = info.ms_update * 1e-3;
  result.xpay = info.ms_xpay * 1e-3;
  return true;
}

void CGHIP::doTransferFrom() { chk(cghip_download_vector(ctx, CGHIP_VEC_X, x), "download x"); }

bool CGHIP::solveFused(int maxIterations, floatType tolerance, FusedResult &result) {
  lastLoop = "hooks";
  if (mixed) return solveMixed(maxIterations, tolerance, result);
  if (!fused) return false;
  lastLoop = graph ? "fused-graph" : "fused";
  cghip_solve_opts o{};
  o.precond = preconditioner == PreconditionerJacobi ? CGHIP_PRECOND_JACOBI : CGHIP_PRECOND_NONE;
  o.max_iter = maxIterations;
  o.tol = tolerance;
  o.batch = 0;
  o.profile = graph ? CGHIP_PROFILE_NONE : CGHIP_PROFILE_KERNELS;
  cghip_solve_info info{};
  chk(cghip_solve(ctx, &o, &info), "solve");
  result.iterations = info.iterations;
  result.residual = info.residual;
  residualHistory(info.iterations, result.residuals);
  result.matvec = (info.ms_matvec + info.ms_init) * 1e-3;
  result.axpy = info.ms_update * 1e-3;
  result.xpay = info.ms_xpay * 1e-3;
  return true;
}

// "backend": {"name": "hip", "device": d, "loop": hooks | fused | fused-graph
// | mixed, "dot": tree | serial, and for mixed solves "mixed": {...}}.
void CGHIP::recordBackend(std::ostream &os) {
  os << "\"name\": \"hip\", \"device\": " << device << ", \"loop\": \"" << lastLoop
     << "\", \"dot\": \"" << (dotMode == CGHIP_DOT_SERIAL ? "serial" : "tree") << "\"";
  if (lastLoop != "mixed") return;
  const auto num = [&os](double v) {
    if (std::isfinite(v))
      os << std::setprecision(17) << v;
    else
      os << "null";
  };
  os << ", \"mixed\": {\"polynomial\": \"" << polyMode << "\"";
  if (learned) {
    os << ", \"method\": \"" << lastLearned.method << "\", \"degree\": " << lastLearned.degree
       << ", \"poly_iterations\": "
       << lastLearned.polyIterations << ", \"jacobi_iterations\": " << lastLearned.jacobiIterations
       << ", \"true_residual\": ";
    num(lastLearned.trueResidual);
  } else {
    os << ", \"inner_iterations\": " << lastMixed.iterations
       << ", \"reliable_updates\": " << lastMixed.updates << ", \"restarts\": " << lastMixed.restarts
       << ", \"true_residual\": ";
    num(lastMixed.true_residual);
    os << ", \"inner_residual\": ";
    num(lastMixed.inner_residual);
  }
  os << "}";
}

CG *CG::getInstance() { return new CGHIP; }
