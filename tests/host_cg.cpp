// Test helper: a host backend of the build's CG driver (cgxx_amd/host/CG.cpp,
// main.cpp), so the driver -- environment parsing, set-up, the reference's
// loop hook by hook, the check, the report and the JSON solve record -- runs
// without a GPU.  The hooks compute what SerialCG's do (SerialCG.cpp:75-161):
// rows and dots added left to right from 0, y += a x, y = x + a y, y = C o x.
// Test infrastructure only; built by tests/test_host_driver.py.
#include <vector>

#include "CG.h"

class HostCG : public CG {
  std::vector<floatType> p, q, r, z;

  floatType *vec(Vector v) {
    switch (v) {
      case VectorK: return k;
      case VectorX: return x;
      case VectorP: return p.data();
      case VectorQ: return q.data();
      case VectorR: return r.data();
      default: return z.data();
    }
  }

  bool supportsMatrixFormat(MatrixFormat) override { return true; }
  bool supportsPreconditioner(Preconditioner p_) override { return p_ == PreconditionerJacobi; }

  void cpy(Vector dst, Vector src) override {
    for (int i = 0; i < N; i++) vec(dst)[i] = vec(src)[i];
  }
  void matvecKernel(Vector in, Vector out) override {
    const floatType *xv = vec(in);
    floatType *yv = vec(out);
    if (matrixFormat == MatrixFormatCRS) {
      for (int i = 0; i < N; i++) {
        floatType t = 0;
        for (int j = matrixCRS->ptr[i]; j < matrixCRS->ptr[i + 1]; j++)
          t += matrixCRS->value[j] * xv[matrixCRS->index[j]];
        yv[i] = t;
      }
    } else if (matrixFormat == MatrixFormatELL) {
      for (int i = 0; i < N; i++) {
        floatType t = 0;
        for (int j = 0; j < matrixELL->length[i]; j++) {
          const int64_t s = (int64_t)j * N + i;
          t += matrixELL->data[s] * xv[matrixELL->index[s]];
        }
        yv[i] = t;
      }
    } else {
      for (int i = 0; i < N; i++) yv[i] = 0;
      for (int e = 0; e < nz; e++) yv[matrixCOO->I[e]] += matrixCOO->V[e] * xv[matrixCOO->J[e]];
    }
  }
  void axpyKernel(floatType a, Vector xi, Vector yi) override {
    const floatType *xv = vec(xi);
    floatType *yv = vec(yi);
    for (int i = 0; i < N; i++) yv[i] += a * xv[i];
  }
  void xpayKernel(Vector xi, floatType a, Vector yi) override {
    const floatType *xv = vec(xi);
    floatType *yv = vec(yi);
    for (int i = 0; i < N; i++) yv[i] = xv[i] + a * yv[i];
  }
  floatType vectorDotKernel(Vector a, Vector b) override {
    const floatType *av = vec(a), *bv = vec(b);
    floatType s = 0;
    for (int i = 0; i < N; i++) s += av[i] * bv[i];
    return s;
  }
  void applyPreconditionerKernel(Vector xi, Vector yi) override {
    const floatType *xv = vec(xi);
    floatType *yv = vec(yi);
    for (int i = 0; i < N; i++) yv[i] = jacobi->C[i] * xv[i];
  }

 public:
  HostCG() : CG(MatrixFormatCRS, PreconditionerJacobi) {}
  void init(const char *matrixFile) override {
    CG::init(matrixFile);
    p.assign(N, 0);
    q.assign(N, 0);
    r.assign(N, 0);
    z.assign(N, 0);
  }
};

CG *CG::getInstance() { return new HostCG; }
