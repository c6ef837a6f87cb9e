// Matrix Market reader and CRS / ELL converters.
//
// Behaviour follows the reference reader (Matrix.cpp:13-111): banner
// "%%MatrixMarket matrix coordinate real general|symmetric" (tokens after
// the first compared case-insensitively), comment and empty lines skipped,
// 1-based indices made 0-based, and for symmetric files every off-diagonal
// entry immediately followed by its mirror, so nz = 2 * nz_file - N and the
// per-row entry order is the file order with mirrors interleaved.  The
// failure messages are the reference's.  Differences, all on inputs the
// reference mishandles: a malformed line, an index out of range or a
// symmetric file with missing diagonal entries is reported as an error
// instead of leaving garbage in the arrays.
//
// The file is read in one piece and parsed in place (strtol/strtod, the
// same correctly rounded decimal conversion as the reference's operator>>),
// which keeps a 100M-entry file at seconds instead of minutes.

#include "Matrix.h"

#include <algorithm>
#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <string>
#include <vector>

namespace {

[[noreturn]] void fatal(const char *msg) {
  std::cerr << msg << std::endl;
  std::exit(1);
}

std::string lower(std::string s) {
  std::transform(s.begin(), s.end(), s.begin(), [](unsigned char c) { return (char)std::tolower(c); });
  return s;
}

// Next line of the first `len` bytes of buf starting at pos (without the
// newline); false at the end.  buf[len] is a terminating NUL.
bool next_line(const std::vector<char> &buf, size_t len, size_t &pos, const char *&b,
               const char *&e) {
  if (pos >= len) return false;
  b = buf.data() + pos;
  const void *nl = std::memchr(b, '\n', len - pos);
  e = nl ? static_cast<const char *>(nl) : buf.data() + len;
  pos = (size_t)(e - buf.data()) + 1;
  return true;
}

}  // namespace

MatrixCOO::MatrixCOO(const char *file) {
  std::vector<char> buf;
  size_t len = 0;
  {
    FILE *f = std::fopen(file, "rb");
    if (!f) fatal("Can't open file with matrix!");
    std::fseek(f, 0, SEEK_END);
    const long size = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    buf.assign(size > 0 ? (size_t)size + 1 : 1, '\0');
    len = size > 0 ? std::fread(buf.data(), 1, (size_t)size, f) : 0;
    std::fclose(f);
    buf[len] = '\0';  // strtol/strtod stop here at the latest
  }
  size_t pos = 0;
  const char *b = nullptr, *e = nullptr;
  std::string head;
  if (next_line(buf, len, pos, b, e)) head.assign(b, e);
  char t[5][256] = {{0}};
  std::sscanf(head.c_str(), "%255s %255s %255s %255s %255s", t[0], t[1], t[2], t[3], t[4]);
  if (std::string(t[0]) != "%%MatrixMarket" || lower(t[1]) != "matrix" ||
      lower(t[2]) != "coordinate" || lower(t[3]) != "real")
    fatal("Only supporting real matrices in coordinate format!");
  const std::string storage = lower(t[4]);
  const bool symmetric = storage == "symmetric";
  if (!symmetric && storage != "general") fatal("Only supporting general or symmetric matrices!");

  do {
    if (!next_line(buf, len, pos, b, e)) fatal("An exception occurred while reading the matrix!");
  } while (b == e || *b == '%');
  long M = 0, NN = 0, NZ = 0;
  {
    std::string dims(b, e);
    if (std::sscanf(dims.c_str(), "%ld %ld %ld", &M, &NN, &NZ) != 3)
      fatal("An exception occurred while reading the matrix!");
  }
  if (NN != M) fatal("Need a square matrix!");
  const long total = symmetric ? 2 * NZ - NN : NZ;
  if (NN <= 0 || NZ < 0 || total < 0 || total > 0x7fffffffL || NN > 0x7fffffffL)
    fatal("An exception occurred while reading the matrix!");
  N = (int)NN;
  nz = (int)total;
  I.reset(new int[nz > 0 ? nz : 1]);
  J.reset(new int[nz > 0 ? nz : 1]);
  V.reset(new floatType[nz > 0 ? nz : 1]);
  nzPerRow.reset(new int[N]());

  int k = 0;
  for (long line = 0; line < NZ; line++) {
    if (!next_line(buf, len, pos, b, e)) fatal("An exception occurred while reading the matrix!");
    // The line ends at e; the buffer continues, so bound the parse by hand.
    char *p = const_cast<char *>(b);
    const long r = std::strtol(p, &p, 10);
    const long c = std::strtol(p, &p, 10);
    const double v = std::strtod(p, &p);
    if (p > e || r < 1 || r > N || c < 1 || c > N)
      fatal("An exception occurred while reading the matrix!");
    if (k >= nz) fatal("An exception occurred while reading the matrix!");
    I[k] = (int)r - 1;
    J[k] = (int)c - 1;
    V[k] = v;
    nzPerRow[I[k]]++;
    k++;
    if (symmetric && r != c) {
      if (k >= nz) fatal("An exception occurred while reading the matrix!");
      I[k] = (int)c - 1;
      J[k] = (int)r - 1;
      V[k] = v;
      nzPerRow[I[k]]++;
      k++;
    }
  }
  if (k != nz) fatal("An exception occurred while reading the matrix!");
}

int MatrixCOO::getMaxNz(int from, int to) const {
  int m = 0;
  for (int i = from; i < to; i++) m = std::max(m, nzPerRow[i]);
  return m;
}

// ---- allocation hooks (Matrix.cpp:147-167) ---------------------------------

void MatrixDataCRS::allocatePtr(int rows) { ptr = new int[rows + 1]; }
void MatrixDataCRS::deallocatePtr() {
  delete[] ptr;
  ptr = nullptr;
}
void MatrixDataCRS::allocateIndexAndValue(int values) {
  index = new int[values > 0 ? values : 1];
  value = new floatType[values > 0 ? values : 1];
}
void MatrixDataCRS::deallocateIndexAndValue() {
  delete[] index;
  delete[] value;
  index = nullptr;
  value = nullptr;
}

void MatrixDataELL::allocateLength(int rows) { length = new int[rows]; }
void MatrixDataELL::deallocateLength() {
  delete[] length;
  length = nullptr;
}
void MatrixDataELL::allocateIndexAndData() {
  index = new int[elements > 0 ? elements : 1]();
  data = new floatType[elements > 0 ? elements : 1]();
}
void MatrixDataELL::deallocateIndexAndData() {
  delete[] index;
  delete[] data;
  index = nullptr;
  data = nullptr;
}

// ---- converters -----------------------------------------------------------

// CRS: exclusive prefix sum of the row counts, then a stable scatter in COO
// order (Matrix.cpp:203-228), so each row keeps the file's entry order.
template <>
void DataMatrix<MatrixDataCRS>::convert(const MatrixCOO &coo) {
  N = coo.N;
  nz = coo.nz;
  allocatePtr(N);
  std::vector<int> cursor(N);
  ptr[0] = 0;
  for (int i = 0; i < N; i++) {
    cursor[i] = ptr[i];
    ptr[i + 1] = ptr[i] + coo.nzPerRow[i];
  }
  allocateIndexAndValue(nz);
  for (int e = 0; e < nz; e++) {
    const int row = coo.I[e];
    index[cursor[row]] = coo.J[e];
    value[cursor[row]] = coo.V[e];
    cursor[row]++;
  }
}

// ELL: width = longest row, slot (j, row) at j * N + row (Matrix.cpp:339-361)
// in 64-bit arithmetic; padding slots are zero (never read by the kernels).
template <>
void DataMatrix<MatrixDataELL>::convert(const MatrixCOO &coo) {
  N = coo.N;
  nz = coo.nz;
  maxNz = coo.getMaxNz();
  elements = (int64_t)N * maxNz;
  allocateLength(N);
  std::memcpy(length, coo.nzPerRow.get(), sizeof(int) * (size_t)N);
  allocateIndexAndData();
  std::vector<int> fill(N, 0);
  for (int e = 0; e < nz; e++) {
    const int row = coo.I[e];
    const int64_t slot = (int64_t)fill[row] * N + row;
    index[slot] = coo.J[e];
    data[slot] = coo.V[e];
    fill[row]++;
  }
}
