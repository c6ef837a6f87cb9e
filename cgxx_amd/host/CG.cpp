// The CG driver: environment, set-up, the PCG loop, the check and the report.
//
// The environment variables, their accepted values and their error messages,
// the progress lines and the report format are the reference's
// (CG.cpp:16-149, 174-270, 387-520), so scripts that drive or parse the
// reference's binaries work unchanged.  The loop is the reference's
// Jacobi-PCG (CG.cpp:284-385), written out below step by step.

#include "CG.h"

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <vector>

namespace {

const char *kMaxIter = "CG_MAX_ITER";
const char *kTolerance = "CG_TOLERANCE";
const char *kCheckTolerance = "CG_CHECK_TOLERANCE";
const char *kMatrixFormat = "CG_MATRIX_FORMAT";
const char *kPreconditioner = "CG_PRECONDITIONER";
const char *kWorkDistribution = "CG_WORK_DISTRIBUTION";
const char *kOverlappedGather = "CG_OVERLAPPED_GATHER";
// Additions of this backend, no reference counterpart.
const char *kMatrixCache = "CG_MATRIX_CACHE";
const char *kRhs = "CG_RHS";

// Uniform double in [0, 1) from SplitMix64 (Steele, Lea & Flood, "Fast
// splittable pseudorandom number generators", OOPSLA 2014): the output for
// state seed after i + 1 steps, so entry i depends on (seed, i) only and
// rows can be filled in any order or in parallel.  The Python mirror is
// cgxx_amd.problems.splitmix_uniform.
double splitmixUniform(unsigned long long seed, unsigned long long i) {
  unsigned long long z = seed + (i + 1) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return (double)(z >> 11) * 0x1.0p-53;
}

const char *env(const char *name) {
  const char *v = std::getenv(name);
  return (v != nullptr && *v != 0) ? v : nullptr;
}

[[noreturn]] void fail(const std::string &msg) {
  std::cerr << msg << std::endl;
  std::exit(1);
}

std::string transformed(const char *s, int (*f)(int)) {
  std::string out(s);
  std::transform(out.begin(), out.end(), out.begin(), [f](unsigned char c) { return (char)f(c); });
  return out;
}

}  // namespace

void CG::parseEnvironment() {
  char *end = nullptr;
  if (const char *v = env(kMaxIter)) {
    errno = 0;
    const long m = std::strtol(v, &end, 0);
    if (errno != 0 || *end != 0 || m <= 0 || m > 0x7fffffffL)
      fail(std::string("Invalid value for ") + kMaxIter + "!");
    maxIterations = (int)m;
  }
  if (const char *v = env(kTolerance)) {
    errno = 0;
    const floatType t = std::strtod(v, &end);
    if (errno != 0 || *end != 0 || !(t > 0)) fail(std::string("Invalid value for ") + kTolerance + "!");
    tolerance = t;
  }
  if (const char *v = env(kCheckTolerance)) {
    errno = 0;
    const floatType t = std::strtod(v, &end);
    // The reference validates `tolerance` here, not the value just read
    // (CG.cpp:67); the accepted values are kept identical.
    if (errno != 0 || *end != 0 || !(tolerance > 0))
      fail(std::string("Invalid value for ") + kCheckTolerance + "!");
    checkTolerance = t;
  }
  if (const char *v = env(kMatrixFormat)) {
    const std::string u = transformed(v, ::toupper);
    if (u == "COO")
      matrixFormat = MatrixFormatCOO;
    else if (u == "CRS")
      matrixFormat = MatrixFormatCRS;
    else if (u == "ELL")
      matrixFormat = MatrixFormatELL;
    else
      fail(std::string("Invalid value for ") + kMatrixFormat + "! (COO, CRS, or ELL)");
    if (!supportsMatrixFormat(matrixFormat)) fail("No support for this matrix format!");
  }
  if (const char *v = env(kPreconditioner)) {
    const std::string l = transformed(v, ::tolower);
    if (l == "none")
      preconditioner = PreconditionerNone;
    else if (l == "jacobi")
      preconditioner = PreconditionerJacobi;
    else
      fail(std::string("Invalid value for ") + kPreconditioner + "! (none, or jacobi)");
    if (preconditioner != PreconditionerNone && !supportsPreconditioner(preconditioner))
      fail("No support for this preconditioner!");
  }
  if (const char *v = env(kWorkDistribution)) {
    const std::string l = transformed(v, ::tolower);
    if (l == "row")
      workDistributionCalc = WorkDistributionByRow;
    else if (l == "nz")
      workDistributionCalc = WorkDistributionByNz;
    else
      fail(std::string("Invalid value for ") + kWorkDistribution + "! (row, or nz)");
  }
  if (const char *v = env(kOverlappedGather)) {
    overlappedGather = std::string(v) != "0";
    if (overlappedGather && (getNumberOfChunks() == -1 || !supportsOverlappedGather()))
      fail("No support for overlapped gather!");
  }
  if (const char *v = env(kMatrixCache)) matrixCache = v;
  if (const char *v = env(kRhs)) {
    const std::string l = transformed(v, ::tolower);
    if (l == "ones") {
      rhsSeed = -1;
    } else if (l.rfind("random:", 0) == 0 && l.size() > 7) {
      errno = 0;
      const long long seed = std::strtoll(l.c_str() + 7, &end, 10);
      if (errno != 0 || *end != 0 || seed < 0)
        fail(std::string("Invalid value for ") + kRhs + "! (ones, or random:<seed>)");
      rhsSeed = seed;
    } else {
      fail(std::string("Invalid value for ") + kRhs + "! (ones, or random:<seed>)");
    }
  }
  hostThreads();  // validates CG_IO_THREADS before any work
}

void CG::allocateMatrixCRS() { matrixCRS.reset(new MatrixCRS); }
void CG::allocateMatrixELL() { matrixELL.reset(new MatrixELL); }
void CG::allocateJacobi() { jacobi.reset(new Jacobi); }
void CG::allocateK() { k = new floatType[N]; }
void CG::deallocateK() {
  delete[] k;
  k = nullptr;
}
void CG::allocateX() { x = new floatType[N]; }
void CG::deallocateX() {
  delete[] x;
  x = nullptr;
}

void CG::init(const char *matrixFile) {
  matrixFileName = matrixFile;
  std::cout << "Reading matrix from " << matrixFile << "..." << std::endl;
  const time_point startIO = now();
  matrixCOO.reset(new MatrixCOO(matrixFile, matrixCache.empty() ? nullptr : matrixCache.c_str()));
  N = matrixCOO->N;
  nz = matrixCOO->nz;

  const time_point startConverting = now();
  switch (matrixFormat) {
  case MatrixFormatCOO:
    break;
  case MatrixFormatCRS:
    std::cout << "Converting matrix to CRS format..." << std::endl;
    allocateMatrixCRS();
    matrixCRS->convert(*matrixCOO);
    break;
  case MatrixFormatELL:
    std::cout << "Converting matrix to ELL format..." << std::endl;
    allocateMatrixELL();
    matrixELL->convert(*matrixCOO);
    break;
  }
  if (preconditioner == PreconditionerJacobi) {
    std::cout << "Initializing Jacobi preconditioner..." << std::endl;
    allocateJacobi();
    jacobi->init(*matrixCOO);
  }
  timing.converting = now() - startConverting;
  timing.io = now() - startIO;

  allocateK();
  if (rhsSeed >= 0) {
    // CG_RHS=random:<seed>: k_i = uniform(-1, 1), the solution unknown
    for (int i = 0; i < N; i++) k[i] = 2.0 * splitmixUniform((unsigned long long)rhsSeed, i) - 1.0;
  } else {
    // k = A * (1, ..., 1)^T, row sums accumulated in COO order (CG.cpp:256-261)
    std::memset(k, 0, sizeof(floatType) * (size_t)N);
    const MatrixCOO &coo = *matrixCOO;
    coo.forEachEntryByRow(hostThreads(), [&](int e) { k[coo.I[e]] += coo.V[e]; });
  }

  allocateX();
  resetX();

  if (matrixFormat != MatrixFormatCOO) matrixCOO.reset();
}

void CG::resetX() { std::memset(x, 0, sizeof(floatType) * (size_t)N); }

void CG::resetNonInitTimings() {
  const Timing kept = timing;
  timing = Timing{};
  timing.io = kept.io;
  timing.converting = kept.converting;
}

void CG::transferTo() {
  const time_point start = now();
  doTransferTo();
  timing.transferTo = now() - start;
}

void CG::transferFrom() {
  const time_point start = now();
  doTransferFrom();
  timing.transferFrom = now() - start;
}

void CG::matvec(Vector in, Vector out) {
  const time_point start = now();
  matvecKernel(in, out);
  timing.matvec += now() - start;
}

void CG::axpy(floatType a, Vector x_, Vector y_) {
  const time_point start = now();
  axpyKernel(a, x_, y_);
  timing.axpy += now() - start;
}

void CG::xpay(Vector x_, floatType a, Vector y_) {
  const time_point start = now();
  xpayKernel(x_, a, y_);
  timing.xpay += now() - start;
}

floatType CG::vectorDot(Vector a, Vector b) {
  const time_point start = now();
  const floatType r = vectorDotKernel(a, b);
  timing.vectorDot += now() - start;
  return r;
}

void CG::applyPreconditioner(Vector x_, Vector y_) {
  const time_point start = now();
  applyPreconditionerKernel(x_, y_);
  timing.preconditioner += now() - start;
}

// Jacobi-preconditioned CG (Hestenes & Stiefel 1952; preconditioned form as in
// the Templates book), stopping on the recursively updated residual
// ||r_i|| / ||r_0|| <= tolerance.
void CG::solve() {
  std::cout << "Solving..." << std::endl;
  const time_point start = now();

  residualHistory.clear();
  FusedResult fused;
  if (solveFused(maxIterations, tolerance, fused)) {
    iteration = fused.iterations;
    residual = fused.residual;
    residualHistory = std::move(fused.residuals);
    timing.matvec += Timing::duration(fused.matvec);
    timing.axpy += Timing::duration(fused.axpy);
    timing.xpay += Timing::duration(fused.xpay);
    timing.vectorDot += Timing::duration(fused.vectorDot);
    timing.preconditioner += Timing::duration(fused.preconditioner);
    timing.solve = now() - start;
    return;
  }

  const bool plain = preconditioner == PreconditionerNone;
  // r = k - A x
  matvec(VectorX, VectorR);
  xpay(VectorK, -1.0, VectorR);
  // p = B r
  if (plain)
    cpy(VectorP, VectorR);
  else
    applyPreconditioner(VectorR, VectorP);
  floatType r2 = vectorDot(VectorR, VectorR);
  const floatType nrm2_0 = std::sqrt(r2);
  floatType rho = plain ? r2 : vectorDot(VectorP, VectorR);
  residualHistory.push_back(1.0);

  for (iteration = 0; iteration < maxIterations; iteration++) {
    matvec(VectorP, VectorQ);                        // q = A p
    const floatType a = rho / vectorDot(VectorP, VectorQ);
    axpy(a, VectorP, VectorX);                       // x += a p
    axpy(-a, VectorQ, VectorR);                      // r -= a q
    r2 = vectorDot(VectorR, VectorR);
    residual = std::sqrt(r2) / nrm2_0;
    residualHistory.push_back(residual);
    if (residual <= tolerance) {
      iteration++;  // this iteration counts
      break;
    }
    const floatType rho_old = rho;
    if (plain) {
      rho = r2;
    } else {
      applyPreconditioner(VectorR, VectorZ);         // z = B r
      rho = vectorDot(VectorR, VectorZ);
    }
    const floatType b = rho / rho_old;
    xpay(plain ? VectorR : VectorZ, b, VectorP);     // p = z + b p
  }
  timing.solve = now() - start;
}

double CG::matvecBytes() const {
  const double n = N, z = nz;
  switch (matrixFormat) {
  case MatrixFormatCRS: return 12.0 * z + 4.0 * (n + 1) + 16.0 * n;
  case MatrixFormatELL: return 12.0 * z + 4.0 * n + 16.0 * n;
  default: return 16.0 * z + 16.0 * n;  // COO: I, J, V, x, y
  }
}

// ||k - A x|| / ||k|| in FP64 with the host's copy of the matrix (x0 = 0, so
// ||k|| = ||r_0||).
floatType CG::hostTrueResidual() const {
  std::vector<floatType> y((size_t)N, 0.0);
  if (matrixFormat == MatrixFormatCRS) {
    for (int i = 0; i < N; i++)
      for (int e = matrixCRS->ptr[i]; e < matrixCRS->ptr[i + 1]; e++)
        y[i] += matrixCRS->value[e] * x[matrixCRS->index[e]];
  } else if (matrixFormat == MatrixFormatELL) {
    for (int i = 0; i < N; i++)
      for (int j = 0; j < matrixELL->length[i]; j++) {
        const long long s = (long long)j * N + i;
        y[i] += matrixELL->data[s] * x[matrixELL->index[s]];
      }
  } else {
    for (int e = 0; e < nz; e++) y[matrixCOO->I[e]] += matrixCOO->V[e] * x[matrixCOO->J[e]];
  }
  double rr = 0, kk = 0;
  for (int i = 0; i < N; i++) {
    const double r = k[i] - y[i];
    rr += r * r;
    kk += k[i] * k[i];
  }
  return kk > 0 ? std::sqrt(rr / kk) : std::sqrt(rr);
}

bool CG::check() {
  std::cout << "Checking solution..." << std::endl;
  const time_point start = now();
  if (rhsSeed >= 0) {
    // CG_RHS=random: the solution is not known; the FP64 true relative
    // residual is judged against CG_CHECK_TOLERANCE instead.
    trueResidual = hostTrueResidual();
    timing.check = now() - start;
    const bool ok = trueResidual <= checkTolerance;
    checkErrors = ok ? 0 : 1;
    std::ostringstream os;
    os << std::scientific << trueResidual;
    std::cout << "True relative residual: " << os.str() << std::endl;
    std::cout << (ok ? "Solution is correct!" : "Solution is incorrect!") << std::endl;
    return ok;
  }
  int errors = 0;
  for (int i = 0; i < N; i++)
    if (std::abs(x[i] - 1.0) > checkTolerance) errors++;
  timing.check = now() - start;
  checkErrors = errors;
  if (errors == 0)
    std::cout << "Solution is correct!" << std::endl;
  else
    std::cout << errors << " elements are incorrect!" << std::endl;
  return errors == 0;
}

void CG::printPadded(const char *label, const std::string &value) {
  std::cout << std::left << std::setw(25) << label << value << std::endl;
}

void CG::printSummary() {
  std::cout << std::endl << "x = [ ";
  for (int i = 0; i < 10 && i < N; i++) std::cout << x[i] << " ";
  std::cout << "]" << std::endl;

  printPadded("Iterations:", std::to_string(iteration));
  std::ostringstream res;
  res << std::scientific << residual;
  printPadded("Residual:", res.str());
  printPadded("# rows / # nonzeros:", std::to_string(N) + " / " + std::to_string(nz));
  const char *formats[] = {"COO", "CRS", "ELL"};
  printPadded("Matrix format:", formats[matrixFormat]);
  printPadded("Preconditioner:", preconditioner == PreconditionerNone ? "None" : "Jacobi");

  std::cout << std::endl;
  printPadded("IO time:", std::to_string(timing.io.count()));
  double total = 0;
  if (matrixFormat != MatrixFormatCOO || preconditioner != PreconditionerNone) {
    printPadded("Converting time:", std::to_string(timing.converting.count()));
    total += timing.converting.count();
  }
  if (needsTransfer()) {
    printPadded("Transfer to time:", std::to_string(timing.transferTo.count()));
    total += timing.transferTo.count();
  }
  printPadded("Solve time:", std::to_string(timing.solve.count()));
  total += timing.solve.count();
  if (needsTransfer()) {
    printPadded("Transfer from time:", std::to_string(timing.transferFrom.count()));
    total += timing.transferFrom.count();
  }
  printPadded("Check time:", std::to_string(timing.check.count()));
  total += timing.check.count();
  printPadded("Total time (excl. IO):", std::to_string(total));

  std::cout << std::endl;
  const double mv = timing.matvec.count();
  printPadded("MatVec time:", std::to_string(mv));
  const double flops = 2.0 * (iteration + 1) * nz;  // + the initial A x0
  printPadded("MatVec GFLOP/s:", std::to_string(flops / 1e9 / mv));
  printPadded("axpy time:", std::to_string(timing.axpy.count()));
  printPadded("xpay time:", std::to_string(timing.xpay.count()));
  printPadded("vectorDot time:", std::to_string(timing.vectorDot.count()));
  if (preconditioner != PreconditionerNone)
    printPadded("Preconditioner time:", std::to_string(timing.preconditioner.count()));
}

namespace {

void jsonString(std::ostream &os, const std::string &v) {
  os << '"';
  for (const unsigned char c : v) {
    if (c == '"' || c == '\\') {
      os << '\\' << c;
    } else if (c < 0x20) {
      char buf[8];
      std::snprintf(buf, sizeof buf, "\\u%04x", c);
      os << buf;
    } else {
      os << c;
    }
  }
  os << '"';
}

// Shortest text that reads back to the same double; null when not finite.
void jsonNumber(std::ostream &os, double v) {
  if (!std::isfinite(v)) {
    os << "null";
    return;
  }
  char buf[32];
  for (int prec = 15; prec <= 17; prec++) {
    std::snprintf(buf, sizeof buf, "%.*g", prec, v);
    if (std::strtod(buf, nullptr) == v) break;
  }
  os << buf;
}

}  // namespace

void CG::recordBackend(std::ostream &os) { os << "\"name\": \"base\""; }

bool CG::writeRecord(const char *path, int repetition) {
  std::ostringstream os;
  const char *formats[] = {"COO", "CRS", "ELL"};
  os << "{\"matrix\": ";
  jsonString(os, matrixFileName);
  os << ", \"rows\": " << N << ", \"nonzeros\": " << nz << ", \"format\": \"" << formats[matrixFormat]
     << "\", \"preconditioner\": \"" << (preconditioner == PreconditionerNone ? "none" : "jacobi")
     << "\", \"tolerance\": ";
  jsonNumber(os, tolerance);
  os << ", \"max_iterations\": " << maxIterations << ", \"repetition\": " << repetition
     << ", \"iterations\": " << iteration << ", \"residual\": ";
  jsonNumber(os, residual);
  os << ", \"check_tolerance\": ";
  jsonNumber(os, checkTolerance);
  os << ", \"incorrect_elements\": " << checkErrors << ", \"rhs\": \""
     << (rhsSeed >= 0 ? "random:" + std::to_string(rhsSeed) : std::string("ones")) << "\"";
  if (rhsSeed >= 0) {
    os << ", \"true_relative_residual\": ";
    jsonNumber(os, trueResidual);
  }
  os << ", \"residual_history\": [";
  for (size_t i = 0; i < residualHistory.size(); i++) {
    if (i) os << ", ";
    jsonNumber(os, residualHistory[i]);
  }
  os << "], \"times\": {";
  const std::pair<const char *, double> times[] = {
      {"io", timing.io.count()},
      {"converting", timing.converting.count()},
      {"transfer_to", timing.transferTo.count()},
      {"solve", timing.solve.count()},
      {"transfer_from", timing.transferFrom.count()},
      {"check", timing.check.count()},
      {"matvec", timing.matvec.count()},
      {"axpy", timing.axpy.count()},
      {"xpay", timing.xpay.count()},
      {"vector_dot", timing.vectorDot.count()},
      {"preconditioner", timing.preconditioner.count()},
  };
  bool first = true;
  for (const auto &t : times) {
    os << (first ? "" : ", ") << '"' << t.first << "\": ";
    jsonNumber(os, t.second);
    first = false;
  }
  os << "}, \"matvec_gflops\": ";
  jsonNumber(os, 2.0 * (iteration + 1) * nz / 1e9 / timing.matvec.count());
  // algorithmic bytes of one SpMV, and their rate over the solve's SpMVs
  // (the initial A x0 included, as in the GFLOP/s figure)
  os << ", \"matvec_bytes\": ";
  jsonNumber(os, matvecBytes());
  os << ", \"matvec_gbps\": ";
  jsonNumber(os, matvecBytes() * (iteration + 1) / 1e9 / timing.matvec.count());
  if (peakGBs() > 0) {
    os << ", \"peak_gbps\": ";
    jsonNumber(os, peakGBs());
    os << ", \"matvec_peak_fraction\": ";
    jsonNumber(os, matvecBytes() * (iteration + 1) / 1e9 / timing.matvec.count() / peakGBs());
  }
  os << ", \"backend\": {";
  recordBackend(os);
  os << "}}\n";
  FILE *f = std::fopen(path, "a");
  if (f == nullptr) return false;
  const std::string line = os.str();
  const bool ok = std::fwrite(line.data(), 1, line.size(), f) == line.size();
  return std::fclose(f) == 0 && ok;
}

void CG::cleanup() {
  deallocateK();
  deallocateX();
  if (matrixCRS) matrixCRS->deallocate();
  if (matrixELL) matrixELL->deallocate();
  if (jacobi) jacobi->deallocateC();
}
