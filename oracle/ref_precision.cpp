// Linked only into the traced reference binary built by oracle/Makefile
// (target ref).  Before main() runs, it raises std::cout to 17 significant
// digits so that the reference's own DEBUG_SOLVE prints (CG.cpp:306-373)
// carry every bit of r2, rho, dot_pq, a and b.  No reference file is edited.
#include <iostream>

namespace {
struct FullPrecision {
  FullPrecision() { std::cout.precision(17); }
} fullPrecision;
}  // namespace
