// Host-side sparse matrix types of the CG path.
//
// Interface mirrors the reference's Matrix.h:14-111 (MatrixCOO read from a
// Matrix Market file, MatrixDataCRS / MatrixDataELL with virtual allocation
// hooks, DataMatrix<Data>::convert) so that a backend written against the
// reference's CG class finds the same names.  The split/partitioned
// multi-device variants (Matrix.h:113-152) are not part of this single-GPU
// build.
#ifndef CGHIP_HOST_MATRIX_H
#define CGHIP_HOST_MATRIX_H

#include <cstdint>
#include <memory>

#include "def.h"

struct Matrix {
  int N = 0;   // dimension
  int nz = 0;  // nonzeros
};

// Coordinate format as read from the file (Matrix.h:22-45).  Symmetric files
// are expanded on reading: each off-diagonal entry is followed by its mirror.
struct MatrixCOO : Matrix {
  std::unique_ptr<int[]> I;
  std::unique_ptr<int[]> J;
  std::unique_ptr<floatType[]> V;
  std::unique_ptr<int[]> nzPerRow;

  MatrixCOO() = delete;
  explicit MatrixCOO(const char *file);

  int getMaxNz() const { return getMaxNz(0, N); }
  int getMaxNz(int from, int to) const;
};

// CRS storage (Matrix.h:48-71).
struct MatrixDataCRS {
  int *ptr = nullptr;
  int *index = nullptr;
  floatType *value = nullptr;

  virtual ~MatrixDataCRS() {}
  virtual void allocatePtr(int rows);
  virtual void deallocatePtr();
  virtual void allocateIndexAndValue(int values);
  virtual void deallocateIndexAndValue();
  void deallocate() {
    deallocatePtr();
    deallocateIndexAndValue();
  }
};

// ELLPACK storage, column-major (Matrix.h:74-103).  `elements` counts slots
// including padding; it is 64-bit here because N * maxNz exceeds int32 for
// heavy-tailed matrices (the reference's int `elements`, Matrix.h:79,
// overflows there).
struct MatrixDataELL {
  int maxNz = 0;
  int64_t elements = 0;
  int *length = nullptr;
  int *index = nullptr;
  floatType *data = nullptr;

  virtual ~MatrixDataELL() {}
  virtual void allocateLength(int rows);
  virtual void deallocateLength();
  virtual void allocateIndexAndData();
  virtual void deallocateIndexAndData();
  void deallocate() {
    deallocateLength();
    deallocateIndexAndData();
  }
};

template <class Data>
struct DataMatrix : Matrix, Data {
  void convert(const MatrixCOO &coo);
};
using MatrixCRS = DataMatrix<MatrixDataCRS>;
using MatrixELL = DataMatrix<MatrixDataELL>;

#endif
