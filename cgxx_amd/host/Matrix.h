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

#include "Parallel.h"
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
  bool rowsSorted = false;  // I non-decreasing (every general file written by rows)

  MatrixCOO() = delete;
  // Reads `file` on hostThreads() threads.  With `cache` (a path), the
  // expanded arrays are loaded from that binary file when it was written for
  // the same source (size and modification time), else the source is parsed
  // and the cache (re)written.
  explicit MatrixCOO(const char *file, const char *cache = nullptr);

  int getMaxNz() const { return getMaxNz(0, N); }
  int getMaxNz(int from, int to) const;

  // Calls f(e) for every entry e, in COO order within each row, on T
  // threads that each own a range of rows, so per-row state needs no
  // locking and per-row results match a sequential pass in COO order.  With
  // sorted rows each thread visits only its own entries; otherwise each
  // scans I in full and skips the rows it does not own.
  template <class F>
  void forEachEntryByRow(int T, F &&f) const {
    if (rowsSorted) {
      parallelFor(T, [&](int t) {
        long long e0 = splitPoint(nz, T, t), e1 = splitPoint(nz, T, t + 1);
        while (e0 > 0 && e0 < nz && I[e0] == I[e0 - 1]) e0++;
        while (e1 > 0 && e1 < nz && I[e1] == I[e1 - 1]) e1++;
        for (long long e = e0; e < e1; e++) f((int)e);
      });
    } else {
      parallelFor(T, [&](int t) {
        const int r0 = (int)splitPoint(N, T, t), r1 = (int)splitPoint(N, T, t + 1);
        for (int e = 0; e < nz; e++) {
          const int r = I[e];
          if (r >= r0 && r < r1) f(e);
        }
      });
    }
  }

 private:
  void parse(const char *file);
  bool loadCache(const char *cache, const char *file);
  void saveCache(const char *cache, const char *file) const;
  bool countRows();
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
