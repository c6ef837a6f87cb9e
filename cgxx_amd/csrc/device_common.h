// Device helpers shared by the HIP kernel files of the CG path: the fixed-
// order reductions, the CRS/ELL row-block SpMV bodies and the grid-stride
// element-group iterators.  Internal; included by kernels*.hip only.
#pragma once

#include "cghip_kernels.h"

#pragma clang fp contract(off)

namespace cghip {
namespace dev {

constexpr int kRounds = kNnzCap / kBlock;  // load rounds per row block

// The SpMV helpers multiply each stored value by xf(x[col]); XIdent is the
// plain SpMV.  A scaling transform gives (c x) products bit for bit, without
// a stored copy of c x.
struct XIdent {
  template <typename T>
  __device__ T operator()(T v) const {
    return v;
  }
};
constexpr int kEllChunk = 8;               // ELL slots in flight per thread
constexpr int kUnroll = 4;                 // element pairs in flight per thread

// ---------------------------------------------------------------------------
// Reductions.

// Lane l accumulates lane l+off for off = 32, 16, ..., 1; lane 0 ends with
// the wave's sum.
__device__ inline double wave_fold(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

// Sum over the 256 threads, returned to every thread: wave sums w0..w3 are
// added left to right.  Ends with a barrier so that `lds` can be reused.
__device__ inline double block_fold(double v, double *lds) {
  v = wave_fold(v);
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = v;
  __syncthreads();
  double r = ((lds[0] + lds[1]) + lds[2]) + lds[3];
  __syncthreads();
  return r;
}

// Thread t's left-to-right sum of partials t, t + 256, ... (G >= 1 of
// them, written by an earlier kernel).  Loads go out 8 at a time, clamped
// to the last partial and dropped by a select past G: a loop of load then
// add waited for every load (up to 24 cache round trips before a vector
// kernel could start streaming).
__device__ inline double thread_partials(const double *partial, int G) {
  double s = 0.0;
  for (int i = threadIdx.x; i < G; i += 8 * kBlock) {
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; u++) v[u] = partial[min(i + u * kBlock, G - 1)];
#pragma unroll
    for (int u = 0; u < 8; u++) s = i + u * kBlock < G ? s + v[u] : s;
  }
  return s;
}

// The same for two arrays (ga, gb >= 1 partials) with both arrays' loads in
// flight together.
__device__ inline void thread_partials2(const double *a, int ga, const double *b, int gb,
                                        double &sa, double &sb) {
  sa = 0.0;
  sb = 0.0;
  const int g = max(ga, gb);
  for (int i = threadIdx.x; i < g; i += 8 * kBlock) {
    double va[8], vb[8];
#pragma unroll
    for (int u = 0; u < 8; u++) {
      va[u] = a[min(i + u * kBlock, ga - 1)];
      vb[u] = b[min(i + u * kBlock, gb - 1)];
    }
#pragma unroll
    for (int u = 0; u < 8; u++) {
      sa = i + u * kBlock < ga ? sa + va[u] : sa;
      sb = i + u * kBlock < gb ? sb + vb[u] : sb;
    }
  }
}

// Fold of G partials written by an earlier kernel: thread t adds partials
// t, t + 256, ... left to right, then the block fold.  Every block that calls
// it gets the same value.
__device__ inline double fold_partials(const double *partial, int G, double *lds) {
  return block_fold(thread_partials(partial, G), lds);
}

// Two such folds, their loads in flight together.
__device__ inline void fold_partials2(const double *a, int ga, const double *b, int gb,
                                      double *lds, double &ra, double &rb) {
  double sa, sb;
  thread_partials2(a, ga, b, gb, sa, sb);
  ra = block_fold(sa, lds);
  rb = block_fold(sb, lds);
}

// ---------------------------------------------------------------------------
// SpMV.

// s + p[j0] + p[j0 + 1] + ... + p[j1 - 1], added left to right, with 8 LDS
// reads in flight (clamped to [j0, j1); the extra reads are dropped by a
// select).
template <typename X>
__device__ inline X lds_sum8(const X *p, int j0, int j1, X s) {
  for (int j = j0; j < j1; j += 8) {
    X w[8];
#pragma unroll
    for (int u = 0; u < 8; u++) w[u] = p[min(j + u, j1 - 1)];
#pragma unroll
    for (int u = 0; u < 8; u++) s = j + u < j1 ? s + w[u] : s;
  }
  return s;
}

// CSR-stream of row block rb with an epilogue: the thread owning row r
// calls pv = pre(r) as soon as it knows its row (before the products, so
// the epilogue's own loads are in flight with the matrix stream), then
// epi(r, s, pv) with the row's sum s (products added left to right from 0)
// and gets back its contribution to a fused dot (0 for threads without a
// row).
//
// The block stages its entry range through LDS (coalesced loads, products
// into prod), one barrier, then each owner adds its row from LDS: a row of
// at most kShortRow products one read at a time, a longer one 8 reads at a
// time (the 8-wide form measured 60 us slower on 7-entry rows at 256^3 and
// 1.3x faster on the heavy-tailed rows).  Measured and not adopted at
// 256^3: each thread loading its own row's entries with no LDS (352 against
// 334 us: strided loads), and straight-line staging with every gather in
// flight at once (347 against 327 us; the gathers of consecutive entries
// spread over the stencil's distant columns).  For short-row matrices
// (A.short_rows) row block rb is rows kBlock * rb .. + kBlock - 1, which
// saves the dependent load of the row-block table (about 1%).
template <typename V, typename X, typename Pre, typename Epi, typename XF = XIdent>
__device__ inline double crs_rowblock_pre_epi(const Crs<V> &A, const X *__restrict__ x, int rb,
                                              X *__restrict__ prod, Pre pre, Epi epi, XF xf = XF()) {
  const int t = threadIdx.x;
  const int r0 = A.short_rows ? rb * kBlock : A.rb_start[rb];
  const int r1 = A.short_rows ? min(r0 + kBlock, A.n) : A.rb_start[rb + 1];
  const int e0 = A.ptr[r0], e1 = A.ptr[r1];
  const int r = r0 + t;
  const bool own = r < r1;
  double dp = 0.0;
  if (e1 - e0 <= kNnzCap) {
    int c[kRounds];
    V v[kRounds];
#pragma unroll
    for (int u = 0; u < kRounds; u++) {
      const int e = e0 + u * kBlock + t;
      if (e < e1) {
        c[u] = A.index[e];
        v[u] = A.value[e];
      }
    }
    int rb0 = 0, re0 = 0;
    decltype(pre(0)) pv{};
    if (own) {
      rb0 = A.ptr[r];
      re0 = A.ptr[r + 1];
      pv = pre(r);
    }
#pragma unroll
    for (int u = 0; u < kRounds; u++) {
      const int e = e0 + u * kBlock + t;
      if (e < e1) prod[e - e0] = X(v[u]) * xf(x[c[u]]);
    }
    __syncthreads();
    if (own) {
      X s = X(0);
      if (re0 - rb0 <= kShortRow) {
        for (int j = rb0; j < re0; j++) s += prod[j - e0];
      } else {
        s = lds_sum8(prod, rb0 - e0, re0 - e0, s);
      }
      dp = epi(r, s, pv);
    }
    __syncthreads();
  } else {
    // One row longer than kNnzCap: chunks through LDS, thread 0 adds in order.
    X s = X(0);
    decltype(pre(0)) pv{};
    if (t == 0) pv = pre(r0);
    for (int c0 = e0; c0 < e1; c0 += kNnzCap) {
      const int ce = min(c0 + kNnzCap, e1);
      for (int e = c0 + t; e < ce; e += kBlock) prod[e - c0] = X(A.value[e]) * xf(x[A.index[e]]);
      __syncthreads();
      if (t == 0) s = lds_sum8(prod, 0, ce - c0, s);
      __syncthreads();
    }
    if (t == 0) dp = epi(r0, s, pv);
  }
  return dp;
}

// The same without early loads: epi(r, s).
template <typename V, typename X, typename Epi>
__device__ inline double crs_rowblock_epi(const Crs<V> &A, const X *__restrict__ x, int rb,
                                          X *__restrict__ prod, Epi epi) {
  return crs_rowblock_pre_epi<V, X>(
      A, x, rb, prod, [](int) { return 0; }, [&](int r, X s, int) { return epi(r, s); });
}

// y = A x over row block rb; with WITH_DOT, returns the owner's x[r] * y[r]
// (in FP64) for the fused <x, A x>.
template <typename V, typename X, bool WITH_DOT>
__device__ inline double crs_rowblock(const Crs<V> &A, const X *__restrict__ x, X *__restrict__ y,
                                      int rb, X *__restrict__ prod) {
  return crs_rowblock_pre_epi<V, X>(
      A, x, rb, prod, [&](int r) { return WITH_DOT ? x[r] : X(0); },
      [&](int r, X s, X xr) -> double {
        y[r] = s;
        return WITH_DOT ? (double)xr * (double)s : 0.0;
      });
}

// ELL row r (column-major slots j * n + r): pv = pre(r) first (the
// epilogue's own loads in flight with the slots), the thread adds its live
// slots left to right from 0 and calls epi(r, s, pv).  The matrix streams
// (length, index, data) are read once, so they are loaded non-temporally
// and leave the caches to the gathered x.  With O32 (allowed when
// width * n * sizeof(V) <= 2^32, see ell_o32) the slots are addressed by
// 32-bit byte offsets from the two array bases, which keeps the FP64
// kernels at 58 VGPRs (8 waves per SIMD) instead of 76 with 64-bit
// addresses per slot.
//
// Slots j0 .. j0 + R - 1 of row r (len > j0), added to s left to right.
// Straight-line code, every slot load and then every gather in flight
// before the first product (loads under per-lane branches had made each
// gather wait for the one before): a slot past the row's length loads the
// row's last live slot again (in bounds; the gather then reads x[r]) and
// its product is dropped by a select (a conditional add let the compiler
// sink the slot's loads under it).  R is uniform across the wave
// (ell_row_pre_epi), so no load is spent on chunks no lane needs.
template <int R, typename V, typename X, bool O32>
__device__ inline X ell_chunk(const Ell<V> &A, const X *__restrict__ x, int r, int len, int j0,
                              X s) {
  const int64_t n = A.n;
  const char *ib = reinterpret_cast<const char *>(A.index);
  const char *vb = reinterpret_cast<const char *>(A.data);
  int c[R];
  V v[R];
  X xv[R];
#pragma unroll
  for (int u = 0; u < R; u++) {
    const int j = min(j0 + u, len - 1);
    if (O32) {
      const uint32_t k = (uint32_t)j * (uint32_t)n + (uint32_t)r;
      c[u] = __builtin_nontemporal_load(reinterpret_cast<const int *>(ib + k * 4u));
      v[u] = __builtin_nontemporal_load(reinterpret_cast<const V *>(vb + k * (uint32_t)sizeof(V)));
    } else {
      const int64_t k = (int64_t)j * n + r;
      c[u] = __builtin_nontemporal_load(A.index + k);
      v[u] = __builtin_nontemporal_load(A.data + k);
    }
  }
#pragma unroll
  for (int u = 0; u < R; u++) xv[u] = x[j0 + u < len ? c[u] : r];
#pragma unroll
  for (int u = 0; u < R; u++) {
    const X p = X(v[u]) * xv[u];
    s = j0 + u < len ? s + p : s;
  }
  return s;
}

template <typename V, typename X, bool O32, typename Pre, typename Epi>
__device__ inline double ell_row_pre_epi(const Ell<V> &A, const X *__restrict__ x, int r, Pre pre,
                                         Epi epi) {
  const int len = __builtin_nontemporal_load(A.length + r);
  const auto pv = pre(r);
  X s = X(0);
  // Chunks of up to kEllChunk slots, as many as the wave's longest row
  // needs (uniform); a lane whose row ends inside a chunk sits it out.
  const int W = __builtin_amdgcn_readfirstlane(__ockl_wfred_max_i32(len));
  for (int j0 = 0; j0 < W; j0 += kEllChunk) {
    if (j0 >= len) continue;
    switch (min(kEllChunk, W - j0)) {
      case 1: s = ell_chunk<1, V, X, O32>(A, x, r, len, j0, s); break;
      case 2: s = ell_chunk<2, V, X, O32>(A, x, r, len, j0, s); break;
      case 3: s = ell_chunk<3, V, X, O32>(A, x, r, len, j0, s); break;
      case 4: s = ell_chunk<4, V, X, O32>(A, x, r, len, j0, s); break;
      case 5: s = ell_chunk<5, V, X, O32>(A, x, r, len, j0, s); break;
      case 6: s = ell_chunk<6, V, X, O32>(A, x, r, len, j0, s); break;
      case 7: s = ell_chunk<7, V, X, O32>(A, x, r, len, j0, s); break;
      default: s = ell_chunk<8, V, X, O32>(A, x, r, len, j0, s); break;
    }
  }
  return epi(r, s, pv);
}

// The same over packed FP32 slots (index | value bits << 32): one 8-byte
// non-temporal load per slot.
template <int R, typename XF>
__device__ inline float ell_chunk(const EllPacked &A, const float *__restrict__ x, int r, int len,
                                  int j0, float s, XF xf) {
  const int64_t n = A.n;
  uint64_t e[R];
  float xv[R];
#pragma unroll
  for (int u = 0; u < R; u++)
    e[u] = __builtin_nontemporal_load(A.slot + ((int64_t)min(j0 + u, len - 1) * n + r));
#pragma unroll
  for (int u = 0; u < R; u++) xv[u] = x[j0 + u < len ? (int)(uint32_t)e[u] : r];
#pragma unroll
  for (int u = 0; u < R; u++) {
    const float p = __uint_as_float((uint32_t)(e[u] >> 32)) * xf(xv[u]);
    s = j0 + u < len ? s + p : s;
  }
  return s;
}

template <typename Pre, typename Epi, typename XF = XIdent>
__device__ inline double ell_row_pre_epi(const EllPacked &A, const float *__restrict__ x, int r,
                                         Pre pre, Epi epi, XF xf = XF()) {
  const int len = __builtin_nontemporal_load(A.length + r);
  const auto pv = pre(r);
  float s = 0.0f;
  const int W = __builtin_amdgcn_readfirstlane(__ockl_wfred_max_i32(len));
  for (int j0 = 0; j0 < W; j0 += kEllChunk) {
    if (j0 >= len) continue;
    switch (min(kEllChunk, W - j0)) {
      case 1: s = ell_chunk<1>(A, x, r, len, j0, s, xf); break;
      case 2: s = ell_chunk<2>(A, x, r, len, j0, s, xf); break;
      case 3: s = ell_chunk<3>(A, x, r, len, j0, s, xf); break;
      case 4: s = ell_chunk<4>(A, x, r, len, j0, s, xf); break;
      case 5: s = ell_chunk<5>(A, x, r, len, j0, s, xf); break;
      case 6: s = ell_chunk<6>(A, x, r, len, j0, s, xf); break;
      case 7: s = ell_chunk<7>(A, x, r, len, j0, s, xf); break;
      default: s = ell_chunk<8>(A, x, r, len, j0, s, xf); break;
    }
  }
  return epi(r, s, pv);
}

// DIA row r: pv = pre(r) first, then the stored entries d = 0, 1, ... in
// ascending diagonal order (the CRS order of a row with ascending columns)
// summed left to right from 0, then epi(r, s, pv).  Values are read once
// (non-temporal); x at r + off[d] is contiguous across a wave.
// FP32 values: one stream per diagonal, an entry stored unless its bits are
// kF32Absent (a row record of the six FP32 values instead measured slower:
// 118 against 111 us).  HALF: the row's FP16 record (half_words(ndiag)
// words: one 12-byte load for the 6 diagonals of a 3D 7-point stencil, else
// one load per word), an entry stored unless its bits are kHalfAbsent;
// converted exactly to FP32.  Both skip absent entries (never a product
// with zero), so both give the row sums of the CRS order.  The record
// instead of one 2-byte stream per diagonal and a mask stream: 256^3
// Clenshaw step 83.2 -> 75.9 us (the absent mark alone, per-diagonal
// streams kept: 78.5 us; the record read as three 4-byte loads: 97.4 us).
//
// SYM (HALF only, Dia::up16 set, kSymPairs diagonal pairs): slot k holds
// upper stream k at r and slot kSymPairs + k the same stream at r - o (row
// r's value on the mirrored lower diagonal), each with x at its column; the
// products are added in the stored order (lower diagonals from the
// farthest, then the upper ones), so the row sums are those of the records.
//
// The loads of one DIA row (dia_row_load) are kept apart from its products
// (dia_row_finish) so that a walk can issue two rows' loads before it uses
// either: dia_row_pre_epi is the two in sequence.
template <typename PV>
struct DiaRowLoads {
  uint32_t v[kDiaMax];
  float xv[kDiaMax];
  PV pv;
};

template <bool HALF, bool SYM = false, typename Pre>
__device__ inline auto dia_row_load(const Dia &A, const float *__restrict__ x, int r, Pre pre) {
  DiaRowLoads<decltype(pre(0))> l;
  l.pv = pre(r);
  const char *xb = reinterpret_cast<const char *>(x);
  // Every load is issued before any is used and none waits for the
  // values: x is read at the column clamped into [0, n) for every
  // diagonal, so the absent mark only selects the products.  Raw value
  // bits first: a conversion next to its load would wait for it.
  const int last = A.n - 1;
  if (SYM) {
    // both loads of a stream unconditional (r - o clamped to 0, its value
    // then replaced by the absent mark); up16 is not streamed
    // non-temporally: rows r + o read it again
    constexpr int Q = kSymPairs;
    const char *ub = reinterpret_cast<const char *>(A.up16);
    const uint32_t nb = ((uint32_t)A.n + (uint32_t)kSymPad) * 2u;
#pragma unroll
    for (int k = 0; k < Q; k++) {
      const int o = A.off[Q + k];
      const int rl = max(r - o, 0);
      const char *sk = ub + (uint32_t)k * nb;
      const uint16_t up = *reinterpret_cast<const uint16_t *>(sk + (uint32_t)r * 2u);
      const uint16_t lo = *reinterpret_cast<const uint16_t *>(sk + (uint32_t)rl * 2u);
      l.v[k] = up;
      l.v[Q + k] = r >= o ? lo : kHalfAbsent;
      l.xv[k] = *reinterpret_cast<const float *>(xb + (uint32_t)min(r + o, last) * 4u);
      l.xv[Q + k] = *reinterpret_cast<const float *>(xb + (uint32_t)rl * 4u);
    }
    return l;
  }
  if (HALF) {
    const int hw = half_words(A.ndiag);
    const uint32_t *rec = reinterpret_cast<const uint32_t *>(A.val16) + (int64_t)r * hw;
    uint32_t w[kDiaMax / 2];
    if (hw == 3) {
      typedef uint32_t u3 __attribute__((ext_vector_type(3)));
      const u3 q = __builtin_nontemporal_load(reinterpret_cast<const u3 *>(rec));
      w[0] = q.x;
      w[1] = q.y;
      w[2] = q.z;
    } else {
#pragma unroll
      for (int k = 0; k < kDiaMax / 2; k++)
        if (k < hw) w[k] = __builtin_nontemporal_load(rec + k);
    }
#pragma unroll
    for (int d = 0; d < kDiaMax; d++)
      if (d < A.ndiag) l.v[d] = (w[d >> 1] >> (16 * (d & 1))) & 0xffffu;
  } else {
    const char *vb = reinterpret_cast<const char *>(A.val);
    const uint32_t nv = (uint32_t)A.n * 4u, rv = (uint32_t)r * 4u;
#pragma unroll
    for (int d = 0; d < kDiaMax; d++)
      if (d < A.ndiag)
        l.v[d] = __builtin_nontemporal_load(
            reinterpret_cast<const uint32_t *>(vb + (rv + (uint32_t)d * nv)));
  }
#pragma unroll
  for (int d = 0; d < kDiaMax; d++)
    if (d < A.ndiag) {
      const int col = min(max(r + A.off[d], 0), last);
      l.xv[d] = *reinterpret_cast<const float *>(xb + (uint32_t)col * 4u);
    }
  return l;
}

template <bool HALF, bool SYM = false, typename L, typename Epi, typename XF>
__device__ inline double dia_row_finish(const Dia &A, int r, const L &l, Epi epi, XF xf) {
  float s = 0.0f;
  if (SYM) {
    constexpr int Q = kSymPairs;
#pragma unroll
    for (int k = Q - 1; k >= 0; k--)
      if (l.v[Q + k] != kHalfAbsent)
        s += (float)__builtin_bit_cast(_Float16, (uint16_t)l.v[Q + k]) * xf(l.xv[Q + k]);
#pragma unroll
    for (int k = 0; k < Q; k++)
      if (l.v[k] != kHalfAbsent)
        s += (float)__builtin_bit_cast(_Float16, (uint16_t)l.v[k]) * xf(l.xv[k]);
    return epi(r, s, l.pv);
  }
#pragma unroll
  for (int d = 0; d < kDiaMax; d++)
    if (d < A.ndiag && (HALF ? l.v[d] != kHalfAbsent : l.v[d] != kF32Absent)) {
      const float a = HALF ? (float)__builtin_bit_cast(_Float16, (uint16_t)l.v[d])
                           : __builtin_bit_cast(float, l.v[d]);
      s += a * xf(l.xv[d]);
    }
  return epi(r, s, l.pv);
}

template <bool HALF, bool SYM = false, typename Pre, typename Epi, typename XF = XIdent>
__device__ inline double dia_row_pre_epi(const Dia &A, const float *__restrict__ x, int r, Pre pre,
                                         Epi epi, XF xf = XF()) {
  const auto l = dia_row_load<HALF, SYM>(A, x, r, pre);
  return dia_row_finish<HALF, SYM>(A, r, l, epi, xf);
}

// ELL rows rb*256 + t: y = A x; with WITH_DOT, returns the owner's
// x[r] * y[r] (in FP64) for the fused <x, A x>.
template <typename V, typename X, bool WITH_DOT, bool O32>
__device__ inline double ell_rowblock(const Ell<V> &A, const X *__restrict__ x, X *__restrict__ y,
                                      int rb) {
  const int r = rb * kBlock + (int)threadIdx.x;
  if (r >= A.n) return 0.0;
  return ell_row_pre_epi<V, X, O32>(
      A, x, r, [&](int i) { return WITH_DOT ? x[i] : X(0); },
      [&](int i, X s, X xi) -> double {
        y[i] = s;
        return WITH_DOT ? (double)xi * (double)s : 0.0;
      });
}

// ---------------------------------------------------------------------------
// Grid-stride iteration over element groups.

struct Pairs {
  int64_t T, full, npair, q;
  __device__ explicit Pairs(int n)
      : T((int64_t)gridDim.x * kBlock), full(n >> 1), npair(((int64_t)n + 1) >> 1),
        q((int64_t)blockIdx.x * kBlock + threadIdx.x) {}
};

__device__ inline const double2 *v2(const double *p) { return reinterpret_cast<const double2 *>(p); }
__device__ inline double2 *v2(double *p) { return reinterpret_cast<double2 *>(p); }

// Pair j of a vector that this kernel only reads, loaded non-temporally:
// read-only streams then do not displace the vectors the kernel writes
// (256^3 FP64: p = z + b p 165 -> 128 us, r -= a q 101 -> 90 us).
__device__ inline double2 v2_stream(const double *p, int64_t j) {
  typedef double d2 __attribute__((ext_vector_type(2)));
  const d2 v = __builtin_nontemporal_load(reinterpret_cast<const d2 *>(p) + j);
  return make_double2(v.x, v.y);
}

// Pair j stored non-temporally: for a vector the next kernels do not read
// (the iterate), so that its stream does not displace the vector the next
// SpMV gathers from the last-level cache.
__device__ inline void v2_store_stream(double *p, int64_t j, double2 v) {
  typedef double d2 __attribute__((ext_vector_type(2)));
  const d2 w = {v.x, v.y};
  __builtin_nontemporal_store(w, reinterpret_cast<d2 *>(p) + j);
}

// The same for quads of a float vector.
__device__ inline float4 v4_stream(const float *p, int64_t j) {
  typedef float f4 __attribute__((ext_vector_type(4)));
  const f4 v = __builtin_nontemporal_load(reinterpret_cast<const f4 *>(p) + j);
  return make_float4(v.x, v.y, v.z, v.w);
}
__device__ inline void v4_store_stream(float *p, int64_t j, float4 v) {
  typedef float f4 __attribute__((ext_vector_type(4)));
  const f4 w = {v.x, v.y, v.z, v.w};
  __builtin_nontemporal_store(w, reinterpret_cast<f4 *>(p) + j);
}

// Quads (4q .. 4q+3) of a float vector: thread gid of T = 256*G visits quads
// gid, gid + T, ... in order, elements in ascending order within a quad.
struct Quads {
  int64_t T, full, nquad, q;
  __device__ explicit Quads(int n)
      : T((int64_t)gridDim.x * kBlock), full(n >> 2), nquad(((int64_t)n + 3) >> 2),
        q((int64_t)blockIdx.x * kBlock + threadIdx.x) {}
};

__device__ inline const float4 *v4(const float *p) { return reinterpret_cast<const float4 *>(p); }
__device__ inline float4 *v4(float *p) { return reinterpret_cast<float4 *>(p); }

}  // namespace dev
}  // namespace cghip
