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
// (SpMV + <p,q> -> MatVec; x/r update + <r,r> + <r,z> -> axpy; Jacobi and
// p update -> xpay; vectorDot and Preconditioner are fused into those and
// report 0).  Environment (in addition to the reference's CG_* variables):
//   CG_HIP_DEVICE   device index (default 0)
//   CG_HIP_FUSED    0: run the reference loop hook by hook instead
//   CG_HIP_DOT      tree (default) | serial: summation order of the dot hook;
//                   serial reproduces SerialCG's left-to-right order
//   CG_HIP_GRAPH    1: replay the fused iterations from a HIP graph (fastest;
//                   the per-category kernel times are then not measured and
//                   print as 0).  Default: one iteration at a time with every
//                   kernel timed by HIP events (one host round trip per
//                   iteration).
#include <cerrno>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <string>

#include "CG.h"
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

  void chk(int rc, const char *what) {
    if (rc != CGHIP_OK) hip_fail(ctx, what);
  }

  bool supportsMatrixFormat(MatrixFormat format) override {
    return format == MatrixFormatCRS || format == MatrixFormatELL;
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
  if (matrixFormat == MatrixFormatCRS) {
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
}

void CGHIP::doTransferFrom() { chk(cghip_download_vector(ctx, CGHIP_VEC_X, x), "download x"); }

bool CGHIP::solveFused(int maxIterations, floatType tolerance, FusedResult &result) {
  if (!fused) return false;
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
  result.matvec = (info.ms_matvec + info.ms_init) * 1e-3;
  result.axpy = info.ms_update * 1e-3;
  result.xpay = info.ms_xpay * 1e-3;
  return true;
}

CG *CG::getInstance() { return new CGHIP; }
