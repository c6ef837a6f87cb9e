// Jacobi preconditioner set-up (behaviour of Preconditioner.cpp:13-23: one
// IEEE division per diagonal entry, in COO order, the last entry of a
// duplicated diagonal winning), on the set-up threads, each owning a range
// of rows.
#include "Preconditioner.h"

#include <cstdlib>
#include <iostream>
#include <vector>

void Jacobi::allocateC(int N) { C = new floatType[N]; }

void Jacobi::deallocateC() {
  delete[] C;
  C = nullptr;
}

void Jacobi::init(const MatrixCOO &coo) {
  allocateC(coo.N);
  std::vector<char> seen(coo.N, 0);
  coo.forEachEntryByRow(hostThreads(), [&](int e) {
    if (coo.I[e] != coo.J[e]) return;
    C[coo.I[e]] = 1 / coo.V[e];
    seen[coo.I[e]] = 1;
  });
  for (int i = 0; i < coo.N; i++) {
    if (!seen[i]) {
      std::cerr << "Row " << i << " has no diagonal entry: no Jacobi preconditioner!"
                << std::endl;
      std::exit(1);
    }
  }
}
