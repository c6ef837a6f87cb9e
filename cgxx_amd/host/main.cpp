// Process entry: `<binary> <matrix.mtx> [<repetitions>]`, as the reference
// (CG.cpp:539-591): parse the environment, read and convert the matrix once,
// then per repetition transfer, solve, transfer back, check and report.
// The exit status is 0 whatever the check says (as in the reference) unless
// CG_STRICT_CHECK is set to a non-zero value, which makes a failed check
// exit with status 2.  With CG_REPORT_JSON=<path>, every repetition also
// appends its figures to <path> as one line of JSON (CG::writeRecord).
#include <cerrno>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <memory>

#include "CG.h"

int main(int argc, char *argv[]) {
  if (argc != 2 && argc != 3) {
    std::cerr << "Usage: " << argv[0] << " <matrix.mtx> [<repetitions>]" << std::endl;
    std::exit(1);
  }
  int repetitions = 1;
  if (argc > 2) {
    char *end = nullptr;
    errno = 0;
    const long r = std::strtol(argv[2], &end, 0);
    if (errno == 0 && *end == 0 && r > 0 && r <= 0x7fffffffL)
      repetitions = (int)r;
    else
      std::cerr << "Invalid value for repetitions! Disabling..." << std::endl;
  }
  const char *strict = std::getenv("CG_STRICT_CHECK");
  const bool strictCheck = strict && *strict && std::strcmp(strict, "0") != 0;
  const char *report = std::getenv("CG_REPORT_JSON");
  if (report != nullptr && *report == 0) report = nullptr;

  std::unique_ptr<CG> cg(CG::getInstance());
  cg->parseEnvironment();
  cg->init(argv[1]);

  bool allCorrect = true;
  for (int i = 0; i < repetitions; i++) {
    if (i != 0) std::cout << std::endl;
    if (cg->needsTransfer()) cg->transferTo();
    cg->solve();
    if (cg->needsTransfer()) cg->transferFrom();
    allCorrect = cg->check() && allCorrect;
    cg->printSummary();
    if (report != nullptr && !cg->writeRecord(report, i)) {
      std::cerr << "Could not write the solve record to " << report << "!" << std::endl;
      std::exit(1);
    }
    cg->resetX();
    cg->resetNonInitTimings();
  }
  cg->cleanup();
  return (strictCheck && !allCorrect) ? 2 : EXIT_SUCCESS;
}
