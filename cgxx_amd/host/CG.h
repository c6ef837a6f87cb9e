// The conjugate-gradient driver: the abstract class a backend plugs into.
//
// Interface of the reference's CG class (CG.h:19-293): the same enums, the
// same protected kernel hooks and allocation hooks, the same environment
// variables, command line and report.  A backend is a subclass that
// implements the hooks and defines CG::getInstance() (one per executable,
// chosen at link time, as in the reference).
//
// One addition: solveFused().  The reference's solve() is not virtual and
// drives the hooks one by one, each synchronous; a device backend that runs
// the whole loop on the device overrides solveFused() and returns true, and
// solve() then takes its iteration count, residual and per-category kernel
// times.  The default returns false and solve() runs the reference loop.
//
// A second addition: writeRecord(), one JSON object per solve with the
// report's figures (CG_REPORT_JSON), extended by the backend through
// recordBackend().
#ifndef CGHIP_HOST_CG_H
#define CGHIP_HOST_CG_H

#include <cassert>
#include <chrono>
#include <memory>
#include <ostream>
#include <string>
#include <vector>

#include "Matrix.h"
#include "Preconditioner.h"
#include "def.h"

class CG {
public:
  enum Vector { VectorK, VectorX, VectorP, VectorQ, VectorR, VectorZ };
  enum MatrixFormat { MatrixFormatCOO, MatrixFormatCRS, MatrixFormatELL };
  enum Preconditioner { PreconditionerNone, PreconditionerJacobi };
  enum WorkDistributionCalc { WorkDistributionByRow, WorkDistributionByNz };

private:
  int iteration = 0;
  int maxIterations = 1000;
  floatType residual = 0;
  floatType tolerance = 1e-9;
  floatType checkTolerance = 1e-5;

  struct Timing {
    using clock = std::chrono::steady_clock;
    using duration = std::chrono::duration<double>;
    duration io{0}, converting{0}, transferTo{0}, transferFrom{0}, check{0};
    duration solve{0}, matvec{0}, axpy{0}, xpay{0}, vectorDot{0}, preconditioner{0};
  };
  Timing timing;
  std::string matrixFileName;
  int checkErrors = -1;  // elements the last check() found incorrect
  // CG_RHS (no reference counterpart): -1 for the reference's k = A * 1,
  // else the seed of k_i = uniform(-1, 1) (randomRhs); check() then judges
  // the FP64 true relative residual instead of x = 1.
  long long rhsSeed = -1;
  floatType trueResidual = -1;  // ||k - A x|| / ||k|| of the last check (random k)
  // Relative residual ||r_i|| / ||r_0|| before the loop and after each
  // iteration of the last solve (the loop's own recursion).
  std::vector<double> residualHistory;

  using time_point = Timing::clock::time_point;
  time_point now() const { return Timing::clock::now(); }

  floatType hostTrueResidual() const;

  // Wall-clock wrappers of the hooks (the hooks are synchronous).
  void matvec(Vector in, Vector out);
  void axpy(floatType a, Vector x, Vector y);
  void xpay(Vector x, floatType a, Vector y);
  floatType vectorDot(Vector a, Vector b);
  void applyPreconditioner(Vector x, Vector y);

protected:
  int N = 0;
  int nz = 0;

  WorkDistributionCalc workDistributionCalc = WorkDistributionByRow;
  bool overlappedGather = false;

  MatrixFormat matrixFormat;
  std::string matrixCache;  // CG_MATRIX_CACHE: binary copy of the read matrix
  std::unique_ptr<MatrixCOO> matrixCOO;
  std::unique_ptr<MatrixCRS> matrixCRS;
  std::unique_ptr<MatrixELL> matrixELL;

  Preconditioner preconditioner;
  std::unique_ptr<Jacobi> jacobi;

  floatType *k = nullptr;  // VectorK, the right-hand side
  floatType *x = nullptr;  // VectorX, the solution

  CG(MatrixFormat defaultMatrixFormat, Preconditioner defaultPreconditioner = PreconditionerNone,
     bool overlappedGather = false)
      : overlappedGather(overlappedGather), matrixFormat(defaultMatrixFormat),
        preconditioner(defaultPreconditioner) {}

  virtual bool supportsMatrixFormat(MatrixFormat format) = 0;
  virtual bool supportsPreconditioner(Preconditioner) { return false; }
  // Multi-device work distribution: -1 = none (the only mode here).
  virtual int getNumberOfChunks() { return -1; }
  virtual bool supportsOverlappedGather() { return false; }

  virtual void allocateMatrixCRS();
  virtual void allocateMatrixELL();
  virtual void allocateJacobi();
  virtual void allocateK();
  virtual void deallocateK();
  virtual void allocateX();
  virtual void deallocateX();

  virtual void doTransferTo() {}
  virtual void doTransferFrom() {}

  // Kernel hooks (CG.h:224-237 of the reference).
  virtual void cpy(Vector dst, Vector src) = 0;                        // dst = src
  virtual void matvecKernel(Vector x, Vector y) = 0;                   // y = A x
  virtual void axpyKernel(floatType a, Vector x, Vector y) = 0;        // y = a x + y
  virtual void xpayKernel(Vector x, floatType a, Vector y) = 0;        // y = x + a y
  virtual floatType vectorDotKernel(Vector a, Vector b) = 0;           // <a, b>
  virtual void applyPreconditionerKernel(Vector, Vector) {             // y = B x
    assert(0 && "Preconditioner not implemented!");
  }

  // Device-resident loop (see the file comment).  Times are seconds.
  struct FusedResult {
    int iterations = 0;
    floatType residual = 0;
    double matvec = 0, axpy = 0, xpay = 0, vectorDot = 0, preconditioner = 0;
    // ||r_i|| / ||r_0|| of the loop's recursion, i = 0 .. iterations (may
    // be left empty or shorter)
    std::vector<double> residuals;
  };
  virtual bool solveFused(int maxIterations, floatType tolerance, FusedResult &result) {
    (void)maxIterations;
    (void)tolerance;
    (void)result;
    return false;
  }

  static void printPadded(const char *label, const std::string &value);

  // Members of the backend's object in the solve record ("backend": {...}),
  // written without the braces; the default writes "name" only.
  virtual void recordBackend(std::ostream &os);
  // Algorithmic bytes of one SpMV in the format the backend runs: values
  // and indices once, x and y once (CRS also the row pointer, ELL the row
  // lengths, COO the row indices).
  virtual double matvecBytes() const;
  // Peak memory bandwidth (GB/s) of the device the SpMV runs on; > 0 adds
  // the SpMV's fraction of it to the solve record.
  virtual double peakGBs() const { return 0; }
  MatrixFormat getMatrixFormat() const { return matrixFormat; }

public:
  virtual ~CG() {}

  virtual void parseEnvironment();
  virtual void init(const char *matrixFile);

  void resetX();
  void resetNonInitTimings();

  virtual bool needsTransfer() { return false; }
  void transferTo();
  void solve();
  void transferFrom();

  virtual void printSummary();
  bool check();
  // Appends one line to `path`: a JSON object with the last solve's
  // figures (those of printSummary, unrounded) and `repetition`.  Returns
  // false if the file cannot be written.
  bool writeRecord(const char *path, int repetition);
  virtual void cleanup();

  int getIterations() const { return iteration; }
  floatType getResidual() const { return residual; }

  static CG *getInstance();
};

#endif
