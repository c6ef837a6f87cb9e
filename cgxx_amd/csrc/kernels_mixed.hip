// Mixed-precision PCG kernels for gfx950: an FP32 inner iteration with FP64
// reliable updates (Sleijpen & van der Vorst, "Reliable updated residuals in
// hybrid Bi-CG methods", Computing 56, 1996; the mixed-precision form of
// Clark et al., "Solving lattice QCD systems of equations using mixed
// precision solvers on GPUs", Comput. Phys. Commun. 181, 2010).
//
// The inner iteration is the fused three-kernel CG of kernels.hip on FP32
// data.  With the Jacobi or polynomial preconditioner it runs on the scaled
// operator A_hat = D^-1/2 A D^-1/2 whose unit diagonal is implicit, so a
// CG step on A_hat is a Jacobi-PCG step on A without any preconditioner
// vector in the loop: the SpMV streams 4 bytes per off-diagonal entry in
// the DIA layout (FP32 value; 2 in the polynomial's Clenshaw steps, which
// read an FP16 copy) or 8 in ELL and CRS (value + int32 index), and the
// vector kernels only y, r, p, q.  Every
// dot product is accumulated in FP64 from the products (double)a * (double)b,
// which are exact for float operands, so only the summation rounds, in a
// fixed tree order.  alpha and beta are computed in FP64 and rounded to FP32
// for the vector updates.  The solution and the true residual live in FP64:
// at each reliable update x64 += s o y32 and r = k - A x64 is recomputed
// with the FP64 matrix; the search direction p32 is kept, so no Krylov
// information is lost as it would be by restarting.  The stop test is the
// FP64 true relative residual.
//
// Vector kernels on FP32 data work on element quads (16-byte loads).

#include "device_common.h"

#pragma clang fp contract(off)

namespace cghip {

namespace {

using namespace dev;

constexpr int kQuadUnroll = 2;  // quads in flight per thread per vector

__global__ __launch_bounds__(kBlock) void k_convert(int64_t n, const double *__restrict__ x,
                                                    float *__restrict__ y) {
  const int64_t T = (int64_t)gridDim.x * kBlock;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += T) y[i] = (float)x[i];
}

__global__ __launch_bounds__(kBlock) void k_fill32(int64_t n, float v, float *__restrict__ y) {
  const int64_t T = (int64_t)gridDim.x * kBlock;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += T) y[i] = v;
}

// s[i] = 1 / sqrt(a_ii); bad[0] = 1 for a row without exactly one diagonal
// entry or with a_ii <= 0 (every writer stores the same value).
__global__ __launch_bounds__(kBlock) void k_scaled_diag(Crs<double> A, double *__restrict__ s,
                                                        int *__restrict__ bad) {
  const int64_t T = (int64_t)gridDim.x * kBlock;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < A.n; i += T) {
    double d = 0.0;
    int cnt = 0;
    for (int e = A.ptr[i]; e < A.ptr[i + 1]; e++)
      if (A.index[e] == (int)i) {
        d = A.value[e];
        cnt++;
      }
    if (cnt != 1 || !(d > 0.0)) {
      bad[0] = 1;
      s[i] = 0.0;
    } else {
      s[i] = 1.0 / sqrt(d);
    }
  }
}

// Off-diagonal entries of row i, in stored order, to offset ptr[i] - i
// (row i has exactly one diagonal entry, checked by k_scaled_diag).
__global__ __launch_bounds__(kBlock) void k_scaled_copy(Crs<double> A, const double *__restrict__ s,
                                                        int *__restrict__ index_nd,
                                                        float *__restrict__ value_nd) {
  const int64_t T = (int64_t)gridDim.x * kBlock;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < A.n; i += T) {
    int k = A.ptr[i] - (int)i;
    const double si = s[i];
    for (int e = A.ptr[i]; e < A.ptr[i + 1]; e++) {
      const int c = A.index[e];
      if (c == (int)i) continue;
      index_nd[k] = c;
      value_nd[k] = (float)((A.value[e] * si) * s[c]);
      k++;
    }
  }
}

__global__ __launch_bounds__(kBlock) void k_scaled_ell(Crs<double> A, const double *__restrict__ s,
                                                       int *__restrict__ length,
                                                       uint64_t *__restrict__ slot) {
  const int64_t T = (int64_t)gridDim.x * kBlock;
  const int64_t n = A.n;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += T) {
    int64_t k = 0;
    const double si = s[i];
    for (int e = A.ptr[i]; e < A.ptr[i + 1]; e++) {
      const int c = A.index[e];
      if (c == (int)i) continue;
      const float v = (float)((A.value[e] * si) * s[c]);
      slot[k * n + i] = (uint64_t)(uint32_t)c | ((uint64_t)__float_as_uint(v) << 32);
      k++;
    }
    length[i] = (int)k;
  }
}

// DIA copy of the scaled off-diagonal entries: entry (i, c) goes to
// diagonal d with D.off[d] == c - i (every off-diagonal offset of the
// matrix is in D.off, checked on the host); row i's slots of val hold
// kF32Absent and its FP16 record kHalfAbsent where no entry is.
__global__ __launch_bounds__(kBlock) void k_scaled_dia(Crs<double> A, const double *__restrict__ s,
                                                       Dia D, float *__restrict__ val,
                                                       uint16_t *__restrict__ val16,
                                                       int *__restrict__ overflow) {
  const int64_t T = (int64_t)gridDim.x * kBlock;
  const int64_t n = A.n;
  const int hw = half_words(D.ndiag);
  const float absent = __builtin_bit_cast(float, kF32Absent);
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += T) {
    const double si = s[i];
    uint16_t *rec = val16 + (int64_t)i * 2 * hw;
    for (int h = 0; h < 2 * hw; h++) rec[h] = kHalfAbsent;
    for (int d = 0; d < D.ndiag; d++) val[(int64_t)d * n + i] = absent;
    for (int e = A.ptr[i]; e < A.ptr[i + 1]; e++) {
      const int c = A.index[e];
      if (c == (int)i) continue;
      const int off = c - (int)i;
      int d = 0;
      while (d < D.ndiag - 1 && D.off[d] != off) d++;
      const float v = (float)((A.value[e] * si) * s[c]);
      if (!(fabsf(v) < 65520.0f)) overflow[0] = 1;  // FP16 would round it to inf
      val[(int64_t)d * n + i] = v;
      rec[d] = __builtin_bit_cast(uint16_t, (_Float16)v);
    }
  }
}

// Symmetric FP16 storage of a mirrored DIA operator (h = ndiag / 2
// diagonal pairs): stream k of up16 is diagonal h + k of the records (the
// entries (i, i + o), o = off[h + k]); row i's value on the mirrored
// diagonal h - 1 - k, the entry (i, i - o), must equal stream k at i - o,
// or be absent where i < o.
__global__ __launch_bounds__(kBlock) void k_dia_upper16(Dia D, uint16_t *__restrict__ up16,
                                                        int *__restrict__ mismatch) {
  const int64_t T = (int64_t)gridDim.x * kBlock;
  const int64_t n = D.n, w = 2 * half_words(D.ndiag);
  const int h = D.ndiag / 2;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += T) {
    const uint16_t *rec = D.val16 + i * w;
    bool bad = false;
    for (int k = 0; k < h; k++) {
      const int64_t o = D.off[h + k];
      up16[k * (n + kSymPad) + i] = rec[h + k];
      const uint16_t mirror = i >= o ? D.val16[(i - o) * w + h + k] : kHalfAbsent;
      bad = bad || rec[h - 1 - k] != mirror;
    }
    if (bad) mismatch[0] = 1;
  }
}

// Row-block loop of the FP32 operator of kind K: for every row i the
// thread owns, pv = pre(i) early, then epi(i, t, pv) with the sum t of the
// stored products of row i (from 0, in stored order); returns the sum of
// the epilogue's dot contributions.
// Chunks of block blockIdx.x: rb0, rb0 + stride, ... below rend.  Round
// robin (block g owns chunks g, g + G, ...), or XCD-contiguous: blocks are
// dispatched to the 8 XCDs round robin, so block g runs on XCD g % 8, and
// that XCD's blocks walk its own contiguous eighth of the chunks; the +-1,
// +-S and +-S^2 neighbours of a stencil row then mostly hit the same L2
// (3.5% faster Clenshaw steps at 256^3).  Needs G % 8 == 0 (else round
// robin).  The DIA kernels walk XCD-contiguously, their dots included (the
// oracle's orc_rowblock_dot_tree_walk restates that assignment); the ELL
// and CRS kernels with a dot walk round robin.
template <bool XCD>
__device__ inline void chunk_walk(int nrb, int &rb0, int &stride, int &rend) {
  rb0 = blockIdx.x, stride = gridDim.x, rend = nrb;
  if (XCD && (gridDim.x & 7) == 0) {
    const int x = blockIdx.x & 7, R = (nrb + 7) >> 3;
    rb0 = x * R + (blockIdx.x >> 3);
    stride = gridDim.x >> 3;
    rend = min(nrb, (x + 1) * R);
  }
}

// Strip walk (DIA steps without a dot, when dia_strips found a plane of Pc
// chunks cut into strips of Wc): the chunks in the order strip-major, then
// plane, then chunk within the strip, and that sequence split into eighths
// per XCD as above.  An XCD's resident blocks then cover a few planes of
// one strip instead of two whole planes, so x[r +- P] is read again a few
// hundred KB of streaming later instead of about one L2's worth.  Each row
// is computed alone, so the walk changes no result.
__device__ inline bool strip_walk(const Dia &d, int &q0, int &stride, int &qend) {
  if (d.strip_chunks <= 0 || (gridDim.x & 7) != 0) return false;
  const int nrb = ell_chunks(d.n), x = blockIdx.x & 7, R = (nrb + 7) >> 3;
  q0 = x * R + (blockIdx.x >> 3);
  stride = gridDim.x >> 3;
  qend = min(nrb, (x + 1) * R);
  return true;
}
__device__ inline int strip_chunk(const Dia &d, int q) {
  const int Pc = d.plane_chunks, Wc = d.strip_chunks;
  const int per = (ell_chunks(d.n) / Pc) * Wc;  // chunks of one strip over all planes
  const int s = q / per, rem = q - s * per;
  const int z = rem / Wc;
  return z * Pc + s * Wc + (rem - z * Wc);
}

template <int K, bool HALF = false, bool XCD = false, bool STRIP = false, int ROWS = 2,
          bool SYM = false, typename Pre, typename Epi, typename XF = XIdent>
__device__ inline double op32_rows(const Op32 &A, const float *__restrict__ x, float *prod, Pre pre,
                                   Epi epi, XF xf = XF()) {
  double tpart = 0.0;
  int q0, qs, qend;
  if (K == kOpDia && STRIP && strip_walk(A.dia, q0, qs, qend)) {
    // ROWS (2 or 4) rows per trip, then two, then one, every row's loads
    // issued before any is used (contributions added in walk order): more
    // bytes in flight per wave for steps that stream little besides the
    // matrix.  256^3 Clenshaw step 76.6 us with one row per trip, 75.4 with
    // two, 74.3 with four; the four-row instances hold 75-95 VGPRs (5-6
    // waves per SIMD), which costs more than it gains when a block walks
    // only a few chunks (128^3: 12% slower), so mp_poly_apply picks them by
    // size.
    int q = q0;
    for (; ROWS >= 4 && q + 3 * qs < qend; q += 4 * qs) {
      const int ra = strip_chunk(A.dia, q) * kBlock + (int)threadIdx.x;
      const int rb = strip_chunk(A.dia, q + qs) * kBlock + (int)threadIdx.x;
      const int rc = strip_chunk(A.dia, q + 2 * qs) * kBlock + (int)threadIdx.x;
      const int rd = strip_chunk(A.dia, q + 3 * qs) * kBlock + (int)threadIdx.x;
      const auto la = dia_row_load<HALF, SYM>(A.dia, x, ra, pre);
      const auto lb = dia_row_load<HALF, SYM>(A.dia, x, rb, pre);
      const auto lc = dia_row_load<HALF, SYM>(A.dia, x, rc, pre);
      const auto ld = dia_row_load<HALF, SYM>(A.dia, x, rd, pre);
      tpart += dia_row_finish<HALF, SYM>(A.dia, ra, la, epi, xf);
      tpart += dia_row_finish<HALF, SYM>(A.dia, rb, lb, epi, xf);
      tpart += dia_row_finish<HALF, SYM>(A.dia, rc, lc, epi, xf);
      tpart += dia_row_finish<HALF, SYM>(A.dia, rd, ld, epi, xf);
    }
    for (; q + qs < qend; q += 2 * qs) {
      const int ra = strip_chunk(A.dia, q) * kBlock + (int)threadIdx.x;
      const int rb = strip_chunk(A.dia, q + qs) * kBlock + (int)threadIdx.x;
      const auto la = dia_row_load<HALF, SYM>(A.dia, x, ra, pre);
      const auto lb = dia_row_load<HALF, SYM>(A.dia, x, rb, pre);
      tpart += dia_row_finish<HALF, SYM>(A.dia, ra, la, epi, xf);
      tpart += dia_row_finish<HALF, SYM>(A.dia, rb, lb, epi, xf);
    }
    if (q < qend) {
      const int r = strip_chunk(A.dia, q) * kBlock + (int)threadIdx.x;
      tpart += dia_row_pre_epi<HALF, SYM>(A.dia, x, r, pre, epi, xf);
    }
  } else if (K == kOpDia) {
    int rb0, stride, rend;
    chunk_walk<XCD>(ell_chunks(A.dia.n), rb0, stride, rend);
    for (int rb = rb0; rb < rend; rb += stride) {
      const int r = rb * kBlock + (int)threadIdx.x;
      if (r < A.dia.n) tpart += dia_row_pre_epi<HALF, SYM>(A.dia, x, r, pre, epi, xf);
    }
  } else if (K == kOpEll) {
    int rb0, stride, rend;
    chunk_walk<XCD>(ell_chunks(A.ell.n), rb0, stride, rend);
    for (int rb = rb0; rb < rend; rb += stride) {
      const int r = rb * kBlock + (int)threadIdx.x;
      if (r < A.ell.n) tpart += ell_row_pre_epi(A.ell, x, r, pre, epi, xf);
    }
  } else {
    for (int rb = blockIdx.x; rb < A.crs.nrb; rb += gridDim.x)
      tpart += crs_rowblock_pre_epi<float, float>(A.crs, x, rb, prod, pre, epi, xf);
  }
  return tpart;
}

// The SGPR cap keeps the DIA instances (8 diagonal offsets in SGPRs) at 8
// blocks of 256 per CU; the CRS and ELL instances use fewer registers.
#define CGHIP_SGPR_CAP __attribute__((amdgpu_num_sgpr(74)))

// x64 += scale o (double)y32 (scale nullable: plain add), over element
// pairs, two pairs in flight per thread (an element at a time ran at
// 5.5 TB/s; four pairs need more than 64 VGPRs).
__global__ __launch_bounds__(kBlock) void k_mp_accum(int n, double *__restrict__ x64,
                                                     const float *__restrict__ y32,
                                                     const double *__restrict__ scale) {
  constexpr int U = 2;
  Pairs P(n);
  const float2 *y2 = reinterpret_cast<const float2 *>(y32);
  for (; P.q + (U - 1) * P.T < P.full; P.q += U * P.T) {
    double2 xv[U], sv[U];
    float2 yv[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int64_t j = P.q + u * P.T;
      xv[u] = v2(x64)[j];
      yv[u] = y2[j];
      sv[u] = scale ? v2_stream(scale, j) : make_double2(0.0, 0.0);
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int64_t j = P.q + u * P.T;
      xv[u].x = xv[u].x + (scale ? sv[u].x * (double)yv[u].x : (double)yv[u].x);
      xv[u].y = xv[u].y + (scale ? sv[u].y * (double)yv[u].y : (double)yv[u].y);
      v2(x64)[j] = xv[u];
    }
  }
  for (; P.q < P.npair; P.q += P.T)
    for (int64_t e = 2 * P.q; e < 2 * P.q + 2 && e < n; e++)
      x64[e] = x64[e] + (scale ? scale[e] * (double)y32[e] : (double)y32[e]);
}

// r64 = k - q64 and the partials of ||r64||^2 (pairs order: a thread's
// pairs in turn, element 2j before 2j + 1), kUnroll pairs in flight.
__global__ __launch_bounds__(kBlock) void k_mp_residual(int n, const double *__restrict__ k,
                                                        const double *__restrict__ q,
                                                        double *__restrict__ r, Partials pa) {
  __shared__ double lds[4];
  Pairs P(n);
  double s = 0.0;
  for (; P.q + (kUnroll - 1) * P.T < P.full; P.q += kUnroll * P.T) {
    double2 kv[kUnroll], qv[kUnroll];
#pragma unroll
    for (int u = 0; u < kUnroll; u++) {
      const int64_t j = P.q + u * P.T;
      kv[u] = v2_stream(k, j);
      qv[u] = v2_stream(q, j);
    }
#pragma unroll
    for (int u = 0; u < kUnroll; u++) {
      const double r0 = kv[u].x - qv[u].x, r1 = kv[u].y - qv[u].y;
      v2(r)[P.q + u * P.T] = make_double2(r0, r1);
      s += r0 * r0;
      s += r1 * r1;
    }
  }
  for (; P.q < P.npair; P.q += P.T) {
    for (int64_t e = 2 * P.q; e < 2 * P.q + 2 && e < n; e++) {
      const double ri = k[e] - q[e];
      r[e] = ri;
      s += ri * ri;
    }
  }
  const double bp = block_fold(s, lds);
  if (threadIdx.x == 0) pa.a[blockIdx.x] = bp;
}

__device__ inline float restart_value(double r, const double *scale, double sv) {
  return scale ? (float)(sv * r) : (float)r;
}

// r32 = (float)(s o r64), y32 = 0, p32 = r32 when first; partials of
// <r32,r32> (quads order: a thread's quads in turn, elements ascending),
// kQuadUnroll quads in flight (an element at a time: 144 us at 256^3).
__global__ __launch_bounds__(kBlock) void k_mp_restart(int n, const double *__restrict__ r64,
                                                       const double *__restrict__ scale,
                                                       float *__restrict__ r32,
                                                       float *__restrict__ y32,
                                                       float *__restrict__ p32, int first,
                                                       Partials pb) {
  __shared__ double lds[4];
  Quads Q(n);
  double s_rr = 0.0;
  for (; Q.q + (kQuadUnroll - 1) * Q.T < Q.full; Q.q += kQuadUnroll * Q.T) {
    double2 ra[kQuadUnroll], rb[kQuadUnroll], sa[kQuadUnroll], sb[kQuadUnroll];
#pragma unroll
    for (int u = 0; u < kQuadUnroll; u++) {
      const int64_t j = Q.q + u * Q.T;
      ra[u] = v2_stream(r64, 2 * j);
      rb[u] = v2_stream(r64, 2 * j + 1);
      if (scale) {
        sa[u] = v2_stream(scale, 2 * j);
        sb[u] = v2_stream(scale, 2 * j + 1);
      } else {
        sa[u] = sb[u] = make_double2(0.0, 0.0);
      }
    }
#pragma unroll
    for (int u = 0; u < kQuadUnroll; u++) {
      const int64_t j = Q.q + u * Q.T;
      const float4 rv = make_float4(restart_value(ra[u].x, scale, sa[u].x),
                                    restart_value(ra[u].y, scale, sa[u].y),
                                    restart_value(rb[u].x, scale, sb[u].x),
                                    restart_value(rb[u].y, scale, sb[u].y));
      v4(r32)[j] = rv;
      v4(y32)[j] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      if (first) v4(p32)[j] = rv;
      s_rr += (double)rv.x * (double)rv.x;
      s_rr += (double)rv.y * (double)rv.y;
      s_rr += (double)rv.z * (double)rv.z;
      s_rr += (double)rv.w * (double)rv.w;
    }
  }
  for (; Q.q < Q.nquad; Q.q += Q.T) {
    for (int64_t e = 4 * Q.q; e < 4 * Q.q + 4 && e < n; e++) {
      const float r = restart_value(r64[e], scale, scale ? scale[e] : 0.0);
      r32[e] = r;
      y32[e] = 0.0f;
      if (first) p32[e] = r;
      s_rr += (double)r * (double)r;
    }
  }
  const double b0 = block_fold(s_rr, lds);
  if (threadIdx.x == 0) pb.a[blockIdx.x] = b0;
}

// One block: the true residual norm, rho and r2 of the restarted inner
// iteration, the stop test on the true residual.
__global__ __launch_bounds__(kBlock) void k_mp_restart_fold(Partials prt, int gt, Partials prr, int gr,
                                                            const double *rzp, int gz, int first,
                                                            MixedState *st, History h) {
  __shared__ double lds[4];
  double rt2, rr;
  fold_partials2(prt.a, gt, prr.a, gr, lds, rt2, rr);
  const double rz = fold_partials(rzp, gz, lds);
  if (threadIdx.x == 0) {
    MixedScalars s = st->next;
    if (first) {
      st->nrm2_0 = sqrt(rt2);
      st->inner0 = sqrt(rr);
      st->updates = 0;
      s.iteration = 0;
    } else {
      st->updates += 1;
    }
    st->true_res = sqrt(rt2) / st->nrm2_0;
    s.rho = rz;
    s.r2 = rr;
    s.r2_ref = rr;
    s.residual = sqrt(rr) / st->inner0;
    s.done = st->true_res <= st->tol ? 1 : 0;
    st->next = s;
    st->cur = s;
    if (first) {
      if (h.r2 && h.cap > 0) h.r2[0] = rr;
      if (h.rho && h.cap > 0) h.rho[0] = rz;
    }
  }
}

__device__ inline bool mp_begin(MixedState *st) {
  const MixedScalars s = st->next;
  const bool stop = s.done != 0 || s.iteration >= st->max_iter;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    MixedScalars c = s;
    if (stop && c.done == 0) c.done = 1;
    st->cur = c;
  }
  return !stop;
}

// q32 = A32 p32 (row sums in stored order in FP32; + p_i with the implicit
// unit diagonal) and partials of <p,q>.  ELL and DIA hold the scaled
// operator only (unit diagonal); CRS also the unscaled FP32 copy.
template <int K>
__global__ __launch_bounds__(kBlock) CGHIP_SGPR_CAP void k_mp_matvec_dot(Op32 A, const float *__restrict__ p,
                                                          float *__restrict__ q, Partials pq,
                                                          MixedState *st) {
  __shared__ float prod[kNnzCap];
  __shared__ double lds[4];
  if (!mp_begin(st)) return;
  const int unit = K == kOpCrs ? A.crs.unit : 1;
  const double tpart = op32_rows<K, false, K == kOpDia>(
      A, p, prod, [&](int i) { return p[i]; },
      [&](int i, float t, float pi) -> double {
        const float qi = unit ? pi + t : t;
        q[i] = qi;
        return (double)pi * (double)qi;
      });
  const double bp = block_fold(tpart, lds);
  if (threadIdx.x == 0) pq.a[blockIdx.x] = bp;
}

__device__ inline void mp_update_elem(float &r, float q, float na, double &s_rr) {
  r = r + na * q;
  s_rr += (double)r * (double)r;
}

// alpha = rho / <p,q> (into st->alpha32 for the p update, which also does
// y32 += alpha p32); r32 -= alpha q32; partials of <r,r>.
__global__ __launch_bounds__(kBlock) void k_mp_update(int n, int gpq, float *__restrict__ r,
                                                      const float *__restrict__ q, Partials ppq,
                                                      Partials prr, MixedState *st, History h) {
  __shared__ double lds[4];
  const MixedScalars c = st->cur;
  if (c.done) return;
  const double pq = fold_partials(ppq.a, gpq, lds);
  const double alpha = c.rho / pq;
  const float na = (float)(-alpha);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->alpha32 = (float)alpha;
    if (h.pq && c.iteration < h.cap) h.pq[c.iteration] = pq;
    if (h.alpha && c.iteration < h.cap) h.alpha[c.iteration] = alpha;
  }
  Quads Q(n);
  double s_rr = 0.0;
  for (; Q.q + (kQuadUnroll - 1) * Q.T < Q.full; Q.q += kQuadUnroll * Q.T) {
    float4 rv[kQuadUnroll], qv[kQuadUnroll];
#pragma unroll
    for (int u = 0; u < kQuadUnroll; u++) {
      const int64_t j = Q.q + u * Q.T;
      rv[u] = v4(r)[j];
      qv[u] = v4(q)[j];
    }
#pragma unroll
    for (int u = 0; u < kQuadUnroll; u++) {
      const int64_t j = Q.q + u * Q.T;
      mp_update_elem(rv[u].x, qv[u].x, na, s_rr);
      mp_update_elem(rv[u].y, qv[u].y, na, s_rr);
      mp_update_elem(rv[u].z, qv[u].z, na, s_rr);
      mp_update_elem(rv[u].w, qv[u].w, na, s_rr);
      v4(r)[j] = rv[u];
    }
  }
  for (; Q.q < Q.nquad; Q.q += Q.T) {
    for (int64_t e = 4 * Q.q; e < 4 * Q.q + 4 && e < n; e++) {
      float re = r[e];
      mp_update_elem(re, q[e], na, s_rr);
      r[e] = re;
    }
  }
  const double b0 = block_fold(s_rr, lds);
  if (threadIdx.x == 0) prr.a[blockIdx.x] = b0;
}

// y32 += a p32 and p32 = u + b p32 over quads (the old p read once); y32,
// read again only by a reliable update, is streamed non-temporally so that
// p32 stays cached for the next SpMV.
__device__ inline void mp_yp_update(int n, float *__restrict__ y, float *__restrict__ p,
                                    const float *__restrict__ u, float a, float b) {
  Quads Q(n);
  for (; Q.q + (kQuadUnroll - 1) * Q.T < Q.full; Q.q += kQuadUnroll * Q.T) {
    float4 yv[kQuadUnroll], pv[kQuadUnroll], uv[kQuadUnroll];
#pragma unroll
    for (int k = 0; k < kQuadUnroll; k++) {
      const int64_t j = Q.q + k * Q.T;
      yv[k] = v4_stream(y, j);
      pv[k] = v4(p)[j];
      uv[k] = v4(u)[j];
    }
#pragma unroll
    for (int k = 0; k < kQuadUnroll; k++) {
      const int64_t j = Q.q + k * Q.T;
      float4 yo, po;
      yo.x = yv[k].x + a * pv[k].x;
      yo.y = yv[k].y + a * pv[k].y;
      yo.z = yv[k].z + a * pv[k].z;
      yo.w = yv[k].w + a * pv[k].w;
      po.x = uv[k].x + b * pv[k].x;
      po.y = uv[k].y + b * pv[k].y;
      po.z = uv[k].z + b * pv[k].z;
      po.w = uv[k].w + b * pv[k].w;
      v4_store_stream(y, j, yo);
      v4(p)[j] = po;
    }
  }
  for (; Q.q < Q.nquad; Q.q += Q.T)
    for (int64_t e = 4 * Q.q; e < 4 * Q.q + 4 && e < n; e++) {
      const float pe = p[e];
      y[e] = y[e] + a * pe;
      p[e] = u[e] + b * pe;
    }
}

// Writes `next`: iteration count, rho = rz, r2 = rr, the inner residual and
// the reliable-update trigger; the history entries.
__device__ inline void mp_advance(MixedState *st, const MixedScalars &c, double rr, double rz,
                                  double beta, History h) {
  const int it = c.iteration;
  MixedScalars s = c;
  s.iteration = it + 1;
  s.rho = rz;
  s.r2 = rr;
  s.residual = sqrt(rr) / st->inner0;
  if (s.residual <= st->tol || rr <= st->delta2 * c.r2_ref) s.done = 2;
  st->next = s;
  if (h.r2 && it + 1 < h.cap) h.r2[it + 1] = rr;
  if (h.rho && it + 1 < h.cap) h.rho[it + 1] = rz;
  if (h.beta && it < h.cap) h.beta[it] = beta;
}

// y32 += alpha p32 (alpha from the r update); beta = <r,r>_new / <r,r>;
// p32 = r32 + beta p32 (always: a reliable update keeps p).  Writes `next`.
__global__ __launch_bounds__(kBlock) void k_mp_xpay(int n, float *__restrict__ y,
                                                    float *__restrict__ p,
                                                    const float *__restrict__ r, Partials prr,
                                                    MixedState *st, History h) {
  __shared__ double lds[4];
  const MixedScalars c = st->cur;
  if (c.done) return;
  const double rr = fold_partials(prr.a, gridDim.x, lds);  // the update ran on this grid
  const double beta = rr / c.rho;
  const float a = st->alpha32;
  if (blockIdx.x == 0 && threadIdx.x == 0) mp_advance(st, c, rr, rr, beta, h);
  mp_yp_update(n, y, p, r, a, (float)beta);
}

// ---------------------------------------------------------------------------
// Polynomial preconditioner (Clenshaw steps fused into FP32 SpMVs).

// Which Clenshaw B the step reads from r (see k_poly_step).
enum { kStepPlain = 0, kStepFirst = 1, kStepSecond = 2 };

// Clenshaw step j with one SpMV t = A_hat B_{j+1}:
//   s = (2/b) t - B_{j+1}
//   j > 0: B_j = (c_j r + 2 s) - B_{j+2}
//   j = 0: z   = (c_0 r + s) - B_2, and the partials of <r, z>
// (B_{j+2} = 0 when it does not exist).  B_k = c_k r is never stored: the
// first step (MODE kStepFirst, j = k - 1) gathers c_k r[col] as its B_{j+1}
// and the second (kStepSecond, j = k - 2) takes c_k r[i] as its B_{j+2},
// the same products a stored B_k would give.  st == nullptr runs the step
// unconditionally (restart); otherwise it is skipped once the iteration is
// off.  HALF (DIA only): the operator values come from the FP16 copy.
// A_hat is the scaled operator: its unit diagonal is implicit (the launcher
// refuses any other).  NT: r and B_{j+2}, read once per row, are loaded
// non-temporally (poly_stream picks it by size).  SYM (with HALF): the
// values come from the upper-diagonal streams Dia::up16, the same values
// and products as the records in 2 instead of 4 bytes per stored pair.  The kernel keeps at most
// 80 SGPRs: above
// that, fewer than 8 blocks of 256 fit on a CU of gfx950 and the grid of
// 2048 blocks runs in two rounds (295 us instead of 200 us per step at
// 256^3).
template <bool FINAL, int K, int MODE, bool HALF, bool NT = false, int ROWS = 2, bool SYM = false>
__global__ __launch_bounds__(kBlock) CGHIP_SGPR_CAP void k_poly_step(Op32 A, const float *__restrict__ b1,
                                                      const float *__restrict__ b2,
                                                      const float *__restrict__ r,
                                                      const PolyParams *poly, int j,
                                                      float *__restrict__ b0,
                                                      float *__restrict__ pcopy, Partials prz,
                                                      const MixedState *st) {
  __shared__ float prod[kNnzCap];
  __shared__ double lds[4];
  if (st && st->cur.done) return;
  const float cj = poly->c[j];
  const float tb = poly->two_over_b;
  const float ck = poly->c[poly->degree];
  const double tpart = op32_rows<K, HALF, K == kOpDia || !FINAL, !FINAL, ROWS, SYM>(
      A, MODE == kStepFirst ? r : b1, prod,
      [&](int i) {
        const float ri = NT ? __builtin_nontemporal_load(r + i) : r[i];
        if (MODE == kStepFirst) return float4{ck * ri, ri, 0.0f, 0.0f};
        if (MODE == kStepSecond) return float4{b1[i], ri, ck * ri, 0.0f};
        return float4{b1[i], ri, b2 ? (NT ? __builtin_nontemporal_load(b2 + i) : b2[i]) : 0.0f, 0.0f};
      },
      [&](int i, float t0, float4 pv) -> double {
        const float bi = pv.x, ri = pv.y, bb = pv.z;
        const float t = bi + t0;
        const float sv = tb * t - bi;
        if (FINAL) {
          const float z = (cj * ri + sv) - bb;
          b0[i] = z;
          if (pcopy) pcopy[i] = z;
          return (double)ri * (double)z;
        }
        b0[i] = (cj * ri + 2.0f * sv) - bb;
        return 0.0;
      },
      [ck](float v) { return MODE == kStepFirst ? ck * v : v; });
  if (FINAL) {
    const double bp = block_fold(tpart, lds);
    if (threadIdx.x == 0) prz.a[blockIdx.x] = bp;
  }
}

// y32 += alpha p32 (alpha from the r update); beta = <r,z>_new / <r,z>;
// p32 = z + beta p32; stop test / reliable-update trigger on the inner
// residual.  Writes `next`.
__global__ __launch_bounds__(kBlock) void k_mp_xpay_poly(int n, float *__restrict__ y,
                                                         float *__restrict__ p,
                                                         const float *__restrict__ z, Partials prr,
                                                         Partials prz, int gz, MixedState *st,
                                                         History h) {
  __shared__ double lds[4];
  const MixedScalars c = st->cur;
  if (c.done) return;
  double rr, rz;
  fold_partials2(prr.a, gridDim.x, prz.a, gz, lds, rr, rz);
  const double beta = rz / c.rho;
  const float a = st->alpha32;
  if (blockIdx.x == 0 && threadIdx.x == 0) mp_advance(st, c, rr, rz, beta, h);
  mp_yp_update(n, y, p, z, a, (float)beta);
}

int row_grid(int n) {
  int64_t g = ((int64_t)n + kBlock - 1) / kBlock;
  if (g > 8192) g = 8192;
  if (g < 1) g = 1;
  return (int)g;
}

}  // namespace

hipError_t convert_f64_f32(int64_t n, const double *x, float *y, hipStream_t s) {
  int64_t g = (n + kBlock - 1) / kBlock;
  if (g > 8192) g = 8192;
  if (g < 1) g = 1;
  hipLaunchKernelGGL(k_convert, dim3((unsigned)g), dim3(kBlock), 0, s, n, x, y);
  return hipGetLastError();
}

hipError_t fill32(int64_t n, float v, float *y, hipStream_t s) {
  int64_t g = (n + kBlock - 1) / kBlock;
  if (g > 8192) g = 8192;
  if (g < 1) g = 1;
  hipLaunchKernelGGL(k_fill32, dim3((unsigned)g), dim3(kBlock), 0, s, n, v, y);
  return hipGetLastError();
}

hipError_t scaled_diag(const Crs<double> &A, double *s, int *bad, hipStream_t st) {
  hipLaunchKernelGGL(k_scaled_diag, dim3(row_grid(A.n)), dim3(kBlock), 0, st, A, s, bad);
  return hipGetLastError();
}

hipError_t scaled_copy(const Crs<double> &A, const double *s, int *index_nd, float *value_nd,
                       hipStream_t st) {
  hipLaunchKernelGGL(k_scaled_copy, dim3(row_grid(A.n)), dim3(kBlock), 0, st, A, s, index_nd,
                     value_nd);
  return hipGetLastError();
}

hipError_t scaled_ell(const Crs<double> &A, const double *s, int *length, uint64_t *slot,
                      hipStream_t st) {
  hipLaunchKernelGGL(k_scaled_ell, dim3(row_grid(A.n)), dim3(kBlock), 0, st, A, s, length, slot);
  return hipGetLastError();
}

hipError_t scaled_dia(const Crs<double> &A, const double *s, const Dia &D, float *val,
                      uint16_t *val16, int *overflow, hipStream_t st) {
  hipLaunchKernelGGL(k_scaled_dia, dim3(row_grid(A.n)), dim3(kBlock), 0, st, A, s, D, val, val16,
                     overflow);
  return hipGetLastError();
}

hipError_t dia_upper16(const Dia &D, uint16_t *up16, int *mismatch, hipStream_t st) {
  if (D.ndiag != 2 * kSymPairs || !dia_mirrored(D)) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_dia_upper16, dim3(row_grid(D.n)), dim3(kBlock), 0, st, D, up16, mismatch);
  return hipGetLastError();
}

hipError_t mp_accumulate(int n, double *x64, const float *y32, const double *scale, hipStream_t s) {
  hipLaunchKernelGGL(k_mp_accum, dim3(vec_grid(n)), dim3(kBlock), 0, s, n, x64, y32, scale);
  return hipGetLastError();
}

hipError_t mp_residual(int n, const double *k, const double *q64, double *r64, Partials pa,
                       hipStream_t s) {
  hipLaunchKernelGGL(k_mp_residual, dim3(vec_grid(n)), dim3(kBlock), 0, s, n, k, q64, r64, pa);
  return hipGetLastError();
}

hipError_t mp_restart(int n, const double *r64, const double *scale, float *r32, float *y32,
                      float *p32, bool first, Partials pb, hipStream_t s) {
  hipLaunchKernelGGL(k_mp_restart, dim3(vec_grid4(n)), dim3(kBlock), 0, s, n, r64, scale, r32, y32,
                     p32, (int)first, pb);
  return hipGetLastError();
}

hipError_t mp_restart_fold(int n, Partials prt, Partials prr, const double *rz, int gz, bool first,
                           MixedState *st, History h, hipStream_t s) {
  hipLaunchKernelGGL(k_mp_restart_fold, dim3(1), dim3(kBlock), 0, s, prt, vec_grid(n), prr,
                     vec_grid4(n), rz, gz, (int)first, st, h);
  return hipGetLastError();
}

template <int K, bool HALF, bool NT, int ROWS = 2, bool SYM = false>
static void poly_apply_t(const Op32 &A, const float *r32, const PolyParams *poly, int degree,
                         float *const w[3], float *pcopy, Partials prz, const MixedState *st,
                         hipStream_t s) {
  const dim3 grid(op32_grid(A)), block(kBlock);
  // B_j in w[j % 3]; B_k (= c_k r) and B_{k+1} (= 0) are not stored
  for (int j = degree - 1; j >= 1; j--) {
    const float *b1 = w[(j + 1) % 3], *b2 = (j + 2 <= degree - 1) ? w[(j + 2) % 3] : nullptr;
    if (j == degree - 1)
      hipLaunchKernelGGL((k_poly_step<false, K, kStepFirst, HALF, NT, ROWS, SYM>), grid, block, 0, s, A, nullptr,
                         nullptr, r32, poly, j, w[j % 3], (float *)nullptr, prz, st);
    else if (j == degree - 2)
      hipLaunchKernelGGL((k_poly_step<false, K, kStepSecond, HALF, NT, ROWS, SYM>), grid, block, 0, s, A, b1, nullptr,
                         r32, poly, j, w[j % 3], (float *)nullptr, prz, st);
    else
      hipLaunchKernelGGL((k_poly_step<false, K, kStepPlain, HALF, NT, ROWS, SYM>), grid, block, 0, s, A, b1, b2, r32,
                         poly, j, w[j % 3], (float *)nullptr, prz, st);
  }
  if (degree == 1)
    hipLaunchKernelGGL((k_poly_step<true, K, kStepFirst, HALF, NT, ROWS, SYM>), grid, block, 0, s, A, nullptr, nullptr,
                       r32, poly, 0, w[0], pcopy, prz, st);
  else if (degree == 2)
    hipLaunchKernelGGL((k_poly_step<true, K, kStepSecond, HALF, NT, ROWS, SYM>), grid, block, 0, s, A, w[1], nullptr,
                       r32, poly, 0, w[0], pcopy, prz, st);
  else
    hipLaunchKernelGGL((k_poly_step<true, K, kStepPlain, HALF, NT, ROWS, SYM>), grid, block, 0, s, A, w[1], w[2], r32,
                       poly, 0, w[0], pcopy, prz, st);
}

// Non-temporal r and B_{j+2} in the Clenshaw steps (FP16 DIA values) above
// 2^24 rows, where the step's vectors no longer stay in the 256 MB
// last-level cache between launches: 320^3 step 202 -> 177 us; at 256^3
// they would cost 83 -> 87 us.
static bool poly_stream(int n) { return n > (1 << 24); }

// Four rows per strip-walk trip when every block walks at least 8 chunks
// (256^3: 16 per block of 4096); below that the two-row instances, with
// fewer registers and more waves, are faster (128^3: 2 chunks per block).
static bool poly_rows4(const Op32 &A) {
  return (int64_t)ell_chunks(A.dia.n) >= 8 * (int64_t)op32_grid(A);
}

// The FP16 DIA steps: records, or the upper-diagonal streams (SYM).
template <bool SYM>
static void poly_apply_half(const Op32 &A, const float *r32, const PolyParams *poly, int degree,
                            float *const w[3], float *pcopy, Partials prz, const MixedState *st,
                            hipStream_t s) {
  const bool nt = SYM ? poly_stream_sym(A.dia.n) : poly_stream(A.dia.n);
  if (nt && poly_rows4(A))
    poly_apply_t<kOpDia, true, true, 4, SYM>(A, r32, poly, degree, w, pcopy, prz, st, s);
  else if (nt)
    poly_apply_t<kOpDia, true, true, 2, SYM>(A, r32, poly, degree, w, pcopy, prz, st, s);
  else if (poly_rows4(A))
    poly_apply_t<kOpDia, true, false, 4, SYM>(A, r32, poly, degree, w, pcopy, prz, st, s);
  else
    poly_apply_t<kOpDia, true, false, 2, SYM>(A, r32, poly, degree, w, pcopy, prz, st, s);
}

hipError_t mp_poly_apply(const Op32 &A, const float *r32, const PolyParams *poly, int degree,
                         float *const w[3], float *pcopy, Partials prz, const MixedState *st,
                         hipStream_t s, bool half) {
  switch (A.kind) {
    case kOpDia:
      if (A.dia.unit != 1) return hipErrorInvalidValue;
      if (half && A.dia.up16)
        poly_apply_half<true>(A, r32, poly, degree, w, pcopy, prz, st, s);
      else if (half)
        poly_apply_half<false>(A, r32, poly, degree, w, pcopy, prz, st, s);
      else if (poly_rows4(A))
        poly_apply_t<kOpDia, false, false, 4>(A, r32, poly, degree, w, pcopy, prz, st, s);
      else
        poly_apply_t<kOpDia, false, false>(A, r32, poly, degree, w, pcopy, prz, st, s);
      break;
    case kOpEll:
      if (A.ell.unit != 1) return hipErrorInvalidValue;
      poly_apply_t<kOpEll, false, false>(A, r32, poly, degree, w, pcopy, prz, st, s);
      break;
    default:
      if (A.crs.unit != 1) return hipErrorInvalidValue;
      poly_apply_t<kOpCrs, false, false>(A, r32, poly, degree, w, pcopy, prz, st, s);
  }
  return hipGetLastError();
}

hipError_t mp_xpay_poly(int n, float *y32, float *p32, const float *z32, Partials prr, Partials prz,
                        int gz, MixedState *st, History h, hipStream_t s) {
  hipLaunchKernelGGL(k_mp_xpay_poly, dim3(vec_grid4(n)), dim3(kBlock), 0, s, n, y32, p32, z32, prr,
                     prz, gz, st, h);
  return hipGetLastError();
}

hipError_t mp_matvec_dot(const Op32 &A, const float *p, float *q, Partials pq, MixedState *st,
                         hipStream_t s) {
  const dim3 grid(op32_grid(A)), block(kBlock);
  switch (A.kind) {
    case kOpDia:
      if (A.dia.unit != 1) return hipErrorInvalidValue;
      hipLaunchKernelGGL(k_mp_matvec_dot<kOpDia>, grid, block, 0, s, A, p, q, pq, st);
      break;
    case kOpEll:
      if (A.ell.unit != 1) return hipErrorInvalidValue;
      hipLaunchKernelGGL(k_mp_matvec_dot<kOpEll>, grid, block, 0, s, A, p, q, pq, st);
      break;
    default:
      hipLaunchKernelGGL(k_mp_matvec_dot<kOpCrs>, grid, block, 0, s, A, p, q, pq, st);
  }
  return hipGetLastError();
}

hipError_t mp_update(int n, int gpq, float *r32, const float *q32, Partials pq, Partials pr,
                     MixedState *st, History h, hipStream_t s) {
  hipLaunchKernelGGL(k_mp_update, dim3(vec_grid4(n)), dim3(kBlock), 0, s, n, gpq, r32, q32, pq, pr,
                     st, h);
  return hipGetLastError();
}

hipError_t mp_xpay(int n, float *y32, float *p32, const float *r32, Partials pr, MixedState *st,
                   History h, hipStream_t s) {
  hipLaunchKernelGGL(k_mp_xpay, dim3(vec_grid4(n)), dim3(kBlock), 0, s, n, y32, p32, r32, pr, st,
                     h);
  return hipGetLastError();
}

}  // namespace cghip
