// Test helper: reads a Matrix Market file with the product's C++ reader and
// converters (cgxx_amd/host/Matrix.cpp, Preconditioner.cpp) and writes the
// CRS, ELL and Jacobi arrays as raw binary for comparison with the oracle.
// Usage: host_dump <matrix.mtx> <out.bin> [<cache>]  (threads: CG_IO_THREADS)
#include <cstdio>
#include <cstdlib>

#include "Matrix.h"
#include "Preconditioner.h"

int main(int argc, char **argv) {
  if (argc != 3 && argc != 4) return 1;
  MatrixCOO coo(argv[1], argc == 4 ? argv[3] : nullptr);
  MatrixCRS crs;
  crs.convert(coo);
  MatrixELL ell;
  ell.convert(coo);
  Jacobi jac;
  jac.init(coo);
  FILE *f = std::fopen(argv[2], "wb");
  const int hdr[3] = {coo.N, coo.nz, ell.maxNz};
  std::fwrite(hdr, sizeof(int), 3, f);
  std::fwrite(crs.ptr, sizeof(int), coo.N + 1, f);
  std::fwrite(crs.index, sizeof(int), coo.nz, f);
  std::fwrite(crs.value, sizeof(double), coo.nz, f);
  std::fwrite(ell.length, sizeof(int), coo.N, f);
  std::fwrite(ell.index, sizeof(int), (size_t)ell.elements, f);
  std::fwrite(ell.data, sizeof(double), (size_t)ell.elements, f);
  std::fwrite(jac.C, sizeof(double), coo.N, f);
  std::fclose(f);
  // explicit release, as CG::cleanup does (the matrices own raw arrays)
  crs.deallocate();
  ell.deallocate();
  jac.deallocateC();
  return 0;
}
