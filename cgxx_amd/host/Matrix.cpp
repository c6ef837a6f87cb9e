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
// The file is read in one piece and parsed in place on hostThreads()
// threads: the data lines are cut into one contiguous piece per thread at
// line ends, a count of lines per piece gives each thread the index of its
// first entry (for symmetric files, a first pass over the indices also
// counts the mirrors before each piece), and every thread writes its
// entries straight into I, J and V.  Values of up to 19 significant digits
// are decoded by exact short cuts (decimal_to_double) where these decide the
// rounding, and by strtod otherwise, so every value is the correctly rounded
// double of its decimal text, as with the reference's operator>>.  The
// converters, the Jacobi diagonal and the right-hand side run on the same
// threads, each thread owning a range of rows.

#include "Matrix.h"

#include <fcntl.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <cctype>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <limits>
#include <string>
#include <vector>

namespace {

const char *kReadError = "An exception occurred while reading the matrix!";

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
bool next_line(const char *buf, size_t len, size_t &pos, const char *&b, const char *&e) {
  if (pos >= len) return false;
  b = buf + pos;
  const void *nl = std::memchr(b, '\n', len - pos);
  e = nl ? static_cast<const char *>(nl) : buf + len;
  pos = (size_t)(e - buf) + 1;
  return true;
}

// The whole file into a new NUL-terminated buffer, read by T threads with
// pread (no zero-filling of the buffer first); false if it cannot be read.
bool read_file(const char *file, int T, std::unique_ptr<char[]> &buf, size_t &len) {
  const int fd = ::open(file, O_RDONLY);
  if (fd < 0) return false;
  struct stat st;
  if (::fstat(fd, &st) != 0) {
    ::close(fd);
    return false;
  }
  len = st.st_size > 0 ? (size_t)st.st_size : 0;
  buf.reset(new char[len + 1]);
  std::atomic<bool> bad{false};
  const int parts = len < (64u << 20) ? 1 : T;
  parallelFor(parts, [&](int t) {
    size_t at = (size_t)splitPoint((long long)len, parts, t);
    const size_t end = (size_t)splitPoint((long long)len, parts, t + 1);
    while (at < end) {
      const ssize_t got = ::pread(fd, buf.get() + at, end - at, (off_t)at);
      if (got <= 0) {
        bad = true;
        return;
      }
      at += (size_t)got;
    }
  });
  ::close(fd);
  buf[len] = '\0';  // strtod stops here at the latest
  return !bad;
}

inline bool blank(char c) { return c == ' ' || c == '\t' || c == '\r' || c == '\v' || c == '\f'; }
inline bool digit(char c) { return c >= '0' && c <= '9'; }

// A decimal integer (optional sign, at most 18 digits) after blanks, within
// [p, e).  Advances p past it.
bool parse_index(const char *&p, const char *e, long long &out) {
  while (p < e && blank(*p)) p++;
  bool neg = false;
  if (p < e && (*p == '+' || *p == '-')) neg = *p++ == '-';
  const char *d = p;
  long long v = 0;
  while (p < e && digit(*p) && p - d < 18) v = 10 * v + (*p++ - '0');
  if (p == d || (p < e && digit(*p))) return false;
  out = neg ? -v : v;
  return true;
}

// Exactly representable powers of ten, in double and in the 64-bit
// significand of x87 long double (5^27 < 2^64).
const double kPow10[23] = {1e0,  1e1,  1e2,  1e3,  1e4,  1e5,  1e6,  1e7,  1e8,  1e9,  1e10, 1e11,
                           1e12, 1e13, 1e14, 1e15, 1e16, 1e17, 1e18, 1e19, 1e20, 1e21, 1e22};
const long double kPow10L[28] = {1e0L,  1e1L,  1e2L,  1e3L,  1e4L,  1e5L,  1e6L,  1e7L,  1e8L,  1e9L,
                                 1e10L, 1e11L, 1e12L, 1e13L, 1e14L, 1e15L, 1e16L, 1e17L, 1e18L, 1e19L,
                                 1e20L, 1e21L, 1e22L, 1e23L, 1e24L, 1e25L, 1e26L, 1e27L};
static_assert(std::numeric_limits<long double>::digits == 64, "x87 extended long double expected");

// m * 10^k (m > 0, at most `sig` <= 19 significant digits) correctly rounded
// to double, for the cases that need no big-number arithmetic; false for the
// others.
//  * Clinger's fast path (W. D. Clinger, "How to read floating point numbers
//    accurately", PLDI 1990): for m < 10^15 < 2^53 and |k| <= 22 both
//    operands are exact doubles and one correctly rounded multiply or divide
//    is the answer.
//  * Otherwise, for |k| <= 27: m < 2^64 and 10^|k| are exact long doubles, so
//    one long-double operation gives r within half a long-double ulp of the
//    exact value x.  The midpoints between neighbouring doubles are long
//    doubles (their 11 low significand bits 0x400).  Unless r is one, the
//    nearest midpoint is at least one long-double ulp from r, so x lies on
//    the same side of every midpoint as r and rounding r to double rounds x
//    correctly.  Products stay within [1e-27, 1e47], normal doubles.
bool decimal_to_double(uint64_t m, int sig, int k, double &v) {
  if (sig <= 15 && k >= -22 && k <= 22) {
    v = k >= 0 ? (double)m * kPow10[k] : (double)m / kPow10[-k];
    return true;
  }
  if (k < -27 || k > 27) return false;
  const long double r = k >= 0 ? (long double)m * kPow10L[k] : (long double)m / kPow10L[-k];
  uint64_t significand;
  std::memcpy(&significand, &r, sizeof significand);
  if ((significand & 0x7ff) == 0x400) return false;
  v = (double)r;
  return true;
}

// A real number within [p, e), advancing p past it.  Decimal text with at
// most 19 significant digits goes through decimal_to_double; anything else
// (more digits, far exponents, midpoints, inf/nan/hex forms) goes to strtod,
// correctly rounded in glibc, which also keeps the accepted syntax that of a
// strtod-based reader.  False when nothing was read or the number runs past
// e.
bool parse_value(const char *&p, const char *e, double &out) {
  const char *q = p;
  while (q < e && blank(*q)) q++;
  bool neg = false;
  if (q < e && (*q == '+' || *q == '-')) neg = *q++ == '-';
  uint64_t m = 0;
  int sig = 0, k = 0;
  bool any = false;
  for (; q < e && digit(*q); q++) {
    any = true;
    if (m == 0 && *q == '0') continue;
    if (++sig <= 19) m = 10 * m + (uint64_t)(*q - '0');
  }
  if (sig > 19) k += sig - 19;
  if (q < e && *q == '.') {
    for (q++; q < e && digit(*q); q++) {
      any = true;
      if (m == 0 && *q == '0') {
        k--;
        continue;
      }
      if (++sig <= 19) {
        m = 10 * m + (uint64_t)(*q - '0');
        k--;
      }
    }
  }
  bool fast = any && sig <= 19;
  if (fast && q < e && (*q == 'e' || *q == 'E')) {
    const char *x = q + 1;
    bool eneg = false;
    if (x < e && (*x == '+' || *x == '-')) eneg = *x++ == '-';
    int ev = 0;
    const char *xd = x;
    while (x < e && digit(*x) && x - xd < 5) ev = 10 * ev + (*x++ - '0');
    if (x == xd || (x < e && digit(*x))) fast = false;
    k += eneg ? -ev : ev;
    q = x;
  }
  double v = 0.0;
  if (fast && (q == e || blank(*q)) && (m == 0 || decimal_to_double(m, sig, k, v))) {
    out = neg ? -v : v;
    p = q;
    return true;
  }
  char *end = nullptr;
  out = std::strtod(p, &end);
  if (end == p || end > e) return false;
  p = end;
  return true;
}

// One data line: row, column (1-based, within [1, N]) and value.
bool parse_entry(const char *b, const char *e, int N, int &r, int &c, double &v) {
  long long rr = 0, cc = 0;
  const char *p = b;
  if (!parse_index(p, e, rr) || !parse_index(p, e, cc) || !parse_value(p, e, v)) return false;
  if (rr < 1 || rr > N || cc < 1 || cc > N) return false;
  r = (int)rr - 1;
  c = (int)cc - 1;
  return true;
}

// Start of the line that contains byte `at` or follows it, within [lo, hi].
size_t line_start_after(const char *buf, size_t lo, size_t hi, size_t at) {
  if (at <= lo) return lo;
  if (at >= hi) return hi;
  const void *nl = std::memchr(buf + at - 1, '\n', hi - (at - 1));
  return nl ? (size_t)(static_cast<const char *>(nl) - buf) + 1 : hi;
}

int env_threads() {
  const char *v = std::getenv("CG_IO_THREADS");
  if (v == nullptr || *v == 0) {
    const unsigned h = std::thread::hardware_concurrency();
    return h == 0 ? 1 : (int)std::min(h, 64u);
  }
  char *end = nullptr;
  const long t = std::strtol(v, &end, 10);
  if (*end != 0 || t < 1 || t > 256) fatal("Invalid value for CG_IO_THREADS!");
  return (int)t;
}

// What identifies the source file of a cache: device, inode, size and
// modification time.
struct Source {
  long long dev = -1, ino = -1;
  long long size = -1;
  long long mtime_ns = -1;
};

Source stat_source(const char *file) {
  struct stat st;
  Source s;
  if (::stat(file, &st) == 0) {
    s.dev = (long long)st.st_dev;
    s.ino = (long long)st.st_ino;
    s.size = (long long)st.st_size;
    s.mtime_ns = (long long)st.st_mtim.tv_sec * 1000000000LL + st.st_mtim.tv_nsec;
  }
  return s;
}

// Cache file layout: this header, then I[nz], J[nz] (int32) and V[nz]
// (FP64), native byte order.
struct CacheHeader {
  char magic[8];
  uint32_t version;
  int32_t N;
  int64_t nz;
  int64_t source_size;
  int64_t source_mtime_ns;
  int64_t source_dev;
  int64_t source_ino;
};
const char kCacheMagic[8] = {'C', 'G', 'C', 'O', 'O', 'v', '2', 0};

}  // namespace

int hostThreads() {
  static const int t = env_threads();
  return t;
}

MatrixCOO::MatrixCOO(const char *file, const char *cache) {
  if (cache != nullptr && loadCache(cache, file)) return;
  parse(file);
  if (cache != nullptr) saveCache(cache, file);
}

void MatrixCOO::parse(const char *file) {
  std::unique_ptr<char[]> buf;
  size_t len = 0;
  if (!read_file(file, hostThreads(), buf, len)) fatal("Can't open file with matrix!");
  const char *data = buf.get();
  size_t pos = 0;
  const char *b = nullptr, *e = nullptr;
  std::string head;
  if (next_line(data, len, pos, b, e)) head.assign(b, e);
  char t[5][256] = {{0}};
  std::sscanf(head.c_str(), "%255s %255s %255s %255s %255s", t[0], t[1], t[2], t[3], t[4]);
  if (std::string(t[0]) != "%%MatrixMarket" || lower(t[1]) != "matrix" ||
      lower(t[2]) != "coordinate" || lower(t[3]) != "real")
    fatal("Only supporting real matrices in coordinate format!");
  const std::string storage = lower(t[4]);
  const bool symmetric = storage == "symmetric";
  if (!symmetric && storage != "general") fatal("Only supporting general or symmetric matrices!");

  do {
    if (!next_line(data, len, pos, b, e)) fatal(kReadError);
  } while (b == e || *b == '%');
  long M = 0, NN = 0, NZ = 0;
  {
    std::string dims(b, e);
    if (std::sscanf(dims.c_str(), "%ld %ld %ld", &M, &NN, &NZ) != 3) fatal(kReadError);
  }
  if (NN != M) fatal("Need a square matrix!");
  const long total = symmetric ? 2 * NZ - NN : NZ;
  if (NN <= 0 || NZ < 0 || total < 0 || total > 0x7fffffffL || NN > 0x7fffffffL) fatal(kReadError);
  N = (int)NN;
  nz = (int)total;
  I.reset(new int[nz > 0 ? nz : 1]);
  J.reset(new int[nz > 0 ? nz : 1]);
  V.reset(new floatType[nz > 0 ? nz : 1]);
  nzPerRow.reset(new int[N]());

  // Pieces of the data lines, one per thread, cut at line starts.
  const int T = std::max(1, std::min(hostThreads(), (int)std::min<long>(NZ, 1L << 20) + 1));
  std::vector<size_t> cut(T + 1);
  for (int p = 0; p <= T; p++)
    cut[p] = line_start_after(data, pos, std::max(pos, len),
                              pos + (size_t)splitPoint((long long)(len > pos ? len - pos : 0), T, p));
  std::vector<long long> lines(T + 1, 0);
  parallelFor(T, [&](int p) {
    const char *s = data + cut[p], *x = data + cut[p + 1];
    long long n = std::count(s, x, '\n');
    if (p == T - 1 && x > s && x[-1] != '\n') n++;  // last line without a newline
    lines[p + 1] = n;
  });
  for (int p = 0; p < T; p++) lines[p + 1] += lines[p];
  if (lines[T] < NZ) fatal(kReadError);

  // For symmetric files, the mirrors written before each piece.
  std::vector<long long> mirrors(T + 1, 0);
  std::atomic<bool> bad{false};
  auto each_line = [&](int p, auto &&fn) {
    const char *s = data + cut[p], *x = data + cut[p + 1];
    for (long long line = lines[p]; line < std::min<long long>(lines[p + 1], NZ); line++) {
      const char *nl = static_cast<const char *>(std::memchr(s, '\n', (size_t)(x - s)));
      const char *le = nl ? nl : x;
      if (!fn(line, s, le)) {
        bad = true;
        return;
      }
      s = le + 1;
    }
  };
  if (symmetric) {
    parallelFor(T, [&](int p) {
      long long off = 0;
      each_line(p, [&](long long, const char *s, const char *le) {
        long long r = 0, c = 0;
        const char *q = s;
        if (!parse_index(q, le, r) || !parse_index(q, le, c)) return false;
        off += r != c;
        return true;
      });
      mirrors[p + 1] = off;
    });
    if (bad) fatal(kReadError);
    for (int p = 0; p < T; p++) mirrors[p + 1] += mirrors[p];
    if (NZ + mirrors[T] != (long long)nz) fatal(kReadError);
  }

  std::atomic<bool> unsorted{false};
  parallelFor(T, [&](int p) {
    long long k = lines[p] + mirrors[p];
    int prev = -1;
    bool sorted = true;
    each_line(p, [&](long long, const char *s, const char *le) {
      int r, c;
      double v;
      if (!parse_entry(s, le, N, r, c, v)) return false;
      I[k] = r;
      J[k] = c;
      V[k] = v;
      __atomic_fetch_add(&nzPerRow[r], 1, __ATOMIC_RELAXED);
      sorted = sorted && r >= prev;
      prev = r;
      k++;
      if (symmetric && r != c) {
        I[k] = c;
        J[k] = r;
        V[k] = v;
        __atomic_fetch_add(&nzPerRow[c], 1, __ATOMIC_RELAXED);
        sorted = sorted && c >= prev;
        prev = c;
        k++;
      }
      return true;
    });
    if (!sorted) unsorted = true;
  });
  if (bad) fatal(kReadError);
  // Pieces are each in order; rows also ascend across their seams.
  bool sorted = !unsorted;
  for (int p = 1; p < T && sorted; p++) {
    const long long s = std::min<long long>(lines[p], NZ) + mirrors[p];
    if (s > 0 && s < nz) sorted = I[s - 1] <= I[s];
  }
  rowsSorted = sorted;
}

// Row counts and sortedness of loaded arrays; false if an index is out of
// range.
bool MatrixCOO::countRows() {
  nzPerRow.reset(new int[N]());
  std::atomic<bool> bad{false}, unsorted{false};
  const int T = hostThreads();
  parallelFor(T, [&](int t) {
    const long long e0 = splitPoint(nz, T, t), e1 = splitPoint(nz, T, t + 1);
    int prev = e0 > 0 ? I[e0 - 1] : -1;
    for (long long e = e0; e < e1; e++) {
      const int r = I[e], c = J[e];
      if (r < 0 || r >= N || c < 0 || c >= N) {
        bad = true;
        return;
      }
      if (r < prev) unsorted = true;
      prev = r;
      __atomic_fetch_add(&nzPerRow[r], 1, __ATOMIC_RELAXED);
    }
  });
  if (bad) return false;
  rowsSorted = !unsorted;
  return true;
}

bool MatrixCOO::loadCache(const char *cache, const char *file) {
  const Source src = stat_source(file);
  if (src.size < 0) fatal("Can't open file with matrix!");
  FILE *f = std::fopen(cache, "rb");
  if (!f) return false;
  CacheHeader h;
  bool ok = std::fread(&h, sizeof h, 1, f) == 1 && std::memcmp(h.magic, kCacheMagic, 8) == 0 &&
            h.version == 2 && h.source_size == src.size && h.source_mtime_ns == src.mtime_ns &&
            h.source_dev == src.dev && h.source_ino == src.ino &&
            h.N > 0 && h.nz >= 0 && h.nz <= 0x7fffffffLL;
  if (ok) {
    struct stat st;
    const long long want = (long long)sizeof h + h.nz * (long long)(2 * sizeof(int) + sizeof(floatType));
    ok = ::fstat(fileno(f), &st) == 0 && (long long)st.st_size == want;
  }
  if (ok) {
    N = h.N;
    nz = (int)h.nz;
    I.reset(new int[nz > 0 ? nz : 1]);
    J.reset(new int[nz > 0 ? nz : 1]);
    V.reset(new floatType[nz > 0 ? nz : 1]);
    ok = std::fread(I.get(), sizeof(int), (size_t)nz, f) == (size_t)nz &&
         std::fread(J.get(), sizeof(int), (size_t)nz, f) == (size_t)nz &&
         std::fread(V.get(), sizeof(floatType), (size_t)nz, f) == (size_t)nz;
  }
  std::fclose(f);
  ok = ok && countRows();  // a cache with indices out of range is not used
  if (!ok) {
    I.reset();
    J.reset();
    V.reset();
    nzPerRow.reset();
    N = nz = 0;
  }
  return ok;
}

void MatrixCOO::saveCache(const char *cache, const char *file) const {
  const Source src = stat_source(file);
  CacheHeader h;
  std::memcpy(h.magic, kCacheMagic, 8);
  h.version = 2;
  h.N = N;
  h.nz = nz;
  h.source_size = src.size;
  h.source_mtime_ns = src.mtime_ns;
  h.source_dev = src.dev;
  h.source_ino = src.ino;
  const std::string tmp = std::string(cache) + ".tmp" + std::to_string((long)::getpid());
  FILE *f = std::fopen(tmp.c_str(), "wb");
  bool ok = f != nullptr;
  if (ok) {
    ok = std::fwrite(&h, sizeof h, 1, f) == 1 &&
         std::fwrite(I.get(), sizeof(int), (size_t)nz, f) == (size_t)nz &&
         std::fwrite(J.get(), sizeof(int), (size_t)nz, f) == (size_t)nz &&
         std::fwrite(V.get(), sizeof(floatType), (size_t)nz, f) == (size_t)nz;
    ok = std::fclose(f) == 0 && ok;
  }
  if (ok) ok = std::rename(tmp.c_str(), cache) == 0;
  if (!ok) {
    std::remove(tmp.c_str());
    std::cerr << "Could not write the matrix cache " << cache << "!" << std::endl;
  }
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
  index = new int[elements > 0 ? elements : 1];
  data = new floatType[elements > 0 ? elements : 1];
}
void MatrixDataELL::deallocateIndexAndData() {
  delete[] index;
  delete[] data;
  index = nullptr;
  data = nullptr;
}

// ---- converters -----------------------------------------------------------

// CRS: exclusive prefix sum of the row counts, then a stable scatter in COO
// order (Matrix.cpp:203-228), so each row keeps the file's entry order.  Each
// thread scatters the rows it owns.
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
  coo.forEachEntryByRow(hostThreads(), [&](int e) {
    const int at = cursor[coo.I[e]]++;
    index[at] = coo.J[e];
    value[at] = coo.V[e];
  });
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
  const int T = hostThreads();
  parallelFor(T, [&](int t) {  // zero the padding slots
    const int r0 = (int)splitPoint(N, T, t), r1 = (int)splitPoint(N, T, t + 1);
    for (int j = 0; j < maxNz; j++)
      for (int row = r0; row < r1; row++)
        if (j >= length[row]) {
          index[(int64_t)j * N + row] = 0;
          data[(int64_t)j * N + row] = 0;
        }
  });
  std::vector<int> fill(N, 0);
  coo.forEachEntryByRow(T, [&](int e) {
    const int row = coo.I[e];
    const int64_t slot = (int64_t)fill[row]++ * N + row;
    index[slot] = coo.J[e];
    data[slot] = coo.V[e];
  });
}
