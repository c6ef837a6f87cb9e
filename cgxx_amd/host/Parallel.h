// Host threads for the set-up around the solver (reading the Matrix Market
// file, the CRS/ELL conversions, the Jacobi diagonal and the right-hand
// side): a thread count and a fork-join loop.  No reference counterpart: the
// reference does this set-up on one thread (Matrix.cpp:13-111, 203-228,
// 339-361; CG.cpp:256-261).
#ifndef CGHIP_HOST_PARALLEL_H
#define CGHIP_HOST_PARALLEL_H

#include <thread>
#include <vector>

// CG_IO_THREADS when set (1..256; anything else is an error with the
// reference's "Invalid value for ...!" message), else the hardware threads,
// at most 64.
int hostThreads();

// Runs f(t) for t = 0 .. T-1, f(0) on the calling thread, and joins.  f must
// neither throw nor exit: workers report failures through flags that the
// caller checks after the join.
template <class F>
void parallelFor(int T, F &&f) {
  std::vector<std::thread> pool;
  if (T > 1) pool.reserve((size_t)T - 1);
  for (int t = 1; t < T; t++) pool.emplace_back([&f, t] { f(t); });
  f(0);
  for (auto &th : pool) th.join();
}

// Start of part t of [0, n) cut into T nearly equal parts.
inline long long splitPoint(long long n, int T, int t) {
  return (long long)((__int128)n * t / T);
}

#endif
