// Internal interface between the C ABI (cghip.hip) and the kernels
// (kernels.hip).  Not installed; the public boundary is include/cghip.h.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace cghip {

// Reduction geometry.  The summation order of every reduction is a function
// of these constants and of the vector length / CRS row blocks only -- never
// of the device -- so results are reproducible run to run and restated by
// the CPU oracle.
constexpr int kBlock = 256;      // threads per block (4 wave64s)
constexpr int kVecGMax = 2048;   // max blocks of grid-stride vector kernels
constexpr int kRowsMax = 256;    // max rows in one CRS row block (one per thread)
constexpr int kNnzCap = 2048;    // max nonzeros staged in LDS per row block
constexpr int kShortRow = 8;     // rows of at most this many entries: SpMV without LDS
constexpr int kSpmvGMax = 2048;  // max blocks of the CRS SpMV and FP32 ELL kernels
constexpr int kEllGMax = 0;      // the FP64 SpMV kernels are not capped: one block per row block
constexpr int kDiaGMax = 4096;   // max blocks of the FP32 DIA kernels (scaled operator)
constexpr int kPartialSlot = 8192;  // entries per partial-sum slot (>= every grid above)
constexpr int kDiaMax = 8;       // max diagonals of the FP32 DIA operator

// Vector kernels work on element pairs: G = ceil(pairs / 256), at most 2048.
inline int vec_grid(int n) {
  const int64_t pairs = ((int64_t)n + 1) / 2;
  int64_t g = (pairs + kBlock - 1) / kBlock;
  if (g > kVecGMax) g = kVecGMax;
  if (g < 1) g = 1;
  return (int)g;
}

// FP32 vector kernels work on element quads: G = ceil(quads / 256), <= 2048.
inline int vec_grid4(int n) {
  const int64_t quads = ((int64_t)n + 3) / 4;
  int64_t g = (quads + kBlock - 1) / kBlock;
  if (g > kVecGMax) g = kVecGMax;
  if (g < 1) g = 1;
  return (int)g;
}

// SpMV kernels: block g handles row blocks g, g + G, g + 2G, ...
inline int spmv_grid(int nrb) {
  int64_t g = nrb;
  if (g > kSpmvGMax) g = kSpmvGMax;
  if (g < 1) g = 1;
  return (int)g;
}

// ELL kernels: 256-row chunks, one row per thread; the FP64 ones one block
// per chunk, the FP32 ones at most 2048 blocks.
__host__ __device__ inline int ell_chunks(int n) { return (int)(((int64_t)n + kBlock - 1) / kBlock); }
// Groups of 256 partials folded by k_fold_groups.
inline int fold_groups(int G) { return (G + kBlock - 1) / kBlock; }
inline int ell32_grid(int n) { return spmv_grid(ell_chunks(n)); }
// DIA kernels: the same chunks, up to 4096 blocks (256^3 Clenshaw step: 83.5 us
// at 4096, 85.2 us at 8192, 86.8 us at 2048 blocks).
inline int dia_grid(int n) {
  int64_t g = ell_chunks(n);
  if (g > kDiaGMax) g = kDiaGMax;
  if (g < 1) g = 1;
  return (int)g;
}

// Scalars of the fused PCG loop (CG.cpp:288-291 plus the loop counter and
// the stop flag).  No kernel reads a field that the same kernel writes:
// the p-update kernel (last of an iteration) writes `next`, the SpMV kernel
// (first of an iteration) copies `next` to `cur`, and the x/r-update and
// p-update kernels read `cur`.
struct Scalars {
  double rho, r2, residual, beta;
  int iteration, done;
};

struct SolveState {
  Scalars cur, next;
  double nrm2_0, tol;  // written once, before the loop
  int max_iter, pad;
  double alpha;  // alpha of the running iteration: written by the r update,
                 // read by the p update (which also does x += alpha p)
};

// Per-iteration history: nullable arrays of `cap` entries each; entry i is
// written only while i < cap, so a run longer than the buffer keeps its
// first cap entries.  cap = min(max_iter + 1, kHistCap): the buffer stays
// at most 5 x 8 MiB whatever CG_MAX_ITER is.
constexpr int kHistCap = 1 << 20;
struct History {
  double *r2, *rho, *pq, *alpha, *beta;
  int cap;
};

// Per-block partial sums of the reductions.  A kernel writes its blocks'
// partials with plain stores; the next kernel on the stream (after the
// kernel-boundary coherence point) folds them in every block, in the fixed
// order, so no cross-block handoff ever happens inside a kernel.
struct Partials {
  double *a, *b;  // kPartialSlot entries each
};

template <typename V>
struct Crs {
  const int *ptr, *index;
  const V *value;
  const int *rb_start;  // row-block starts, nrb + 1 entries
  int n, nrb;
  int unit;  // 1: implicit unit diagonal, y_i = x_i + (sum of the stored products)
  // 1: no row has more than kShortRow entries, so (kShortRow * kRowsMax <=
  // kNnzCap) row block rb is exactly rows kRowsMax * rb .. + kRowsMax - 1
  int short_rows;
};
static_assert(kShortRow * kRowsMax <= kNnzCap && kRowsMax == kBlock,
              "short-row blocks are full kRowsMax-row blocks, one row per thread");

template <typename V>
struct Ell {
  const int *length, *index;
  const V *data;
  int n, width;
  int unit;  // 1: implicit unit diagonal, y_i = x_i + (sum of the stored products)
};

// Row blocks of the CRS kernels: greedy, at most kRowsMax rows and at most
// kNnzCap nonzeros, except that a block always takes its first row (a row
// longer than kNnzCap is then streamed through LDS in chunks).
int crs_rowblocks(int n, const int *ptr, int *starts);
// 1 if every row of ptr has at most kShortRow entries and the row blocks
// from crs_rowblocks are the full kRowsMax-row blocks that implies.
int crs_short_rows(int n, const int *ptr, int nrb);

// Message returned by cghip_last_error(NULL) (errors without a context).
void set_global_error(const char *msg);
// Looks up the allocation-free entry point's static scratch array on
// `device` (the current device); cghip_create calls it once per context.
hipError_t noalloc_prepare(int device);

// ---- launches (all on `s`; each returns the launch status) ----------------
hipError_t matvec_crs(const Crs<double> &A, const double *x, double *y, hipStream_t s);
hipError_t matvec_ell(const Ell<double> &A, const double *x, double *y, hipStream_t s);
hipError_t axpy(int n, double a, const double *x, double *y, hipStream_t s);
hipError_t xpay(int n, const double *x, double a, double *y, hipStream_t s);
hipError_t jacobi(int n, const double *C, const double *x, double *y, hipStream_t s);
hipError_t fill(int n, double v, double *y, hipStream_t s);
hipError_t dot_serial(int n, const double *a, const double *b, double *out, hipStream_t s);
// Tree-ordered dot: partials kernel + one-block fold into out[0].
hipError_t dot_tree(int n, const double *a, const double *b, Partials pa, double *out,
                    hipStream_t s);
// out[0] = sum (b - q)^2, out[1] = sum b^2 (vector-kernel tree order)
hipError_t residual_norms(int n, const double *b, const double *q, Partials pa, double *out,
                          hipStream_t s);

// Fused PCG loop kernels (one iteration = matvec_dot, update, xpay).
hipError_t pcg_init(int n, const double *k, double *r, double *p, const double *C, Partials pa,
                    SolveState *st, History h, hipStream_t s);
// q = A p with one partial of <p,q> per row block (CRS) or 256-row chunk
// (ELL) in pq.a, then (pq2 non-null) those partials folded in groups of
// 256 into pq2[0 .. fold_groups(G)), G the number of row blocks / chunks.
// With G <= 256 the loop passes pq2 = null and the update folds the G
// partials itself (pcg_update with gpq = -G).
hipError_t pcg_matvec_dot_crs(const Crs<double> &A, const double *p, double *q, Partials pq,
                              double *pq2, SolveState *st, hipStream_t s);
hipError_t pcg_matvec_dot_ell(const Ell<double> &A, const double *p, double *q, Partials pq,
                              double *pq2, SolveState *st, hipStream_t s);
// alpha = rho / <p,q>; r -= alpha q; partials of <r,r> and <r,Cr>.  <p,q>
// from gpq > 0 group sums in pq.a, or (gpq = -G, G <= 256) from the G
// block partials themselves, folded to the same bits as their one group.
hipError_t pcg_update(int n, int gpq, double *r, const double *q, const double *C, Partials pq,
                      Partials pr, SolveState *st, History h, hipStream_t s);
// x += alpha p (the old p: each p element is read once outside the SpMV);
// stop test; p = C r + beta p unless converged
hipError_t pcg_xpay(int n, double *x, double *p, const double *r, const double *C, Partials pr,
                    SolveState *st, History h, hipStream_t s);

// ---- mixed precision: FP32 inner PCG with FP64 reliable updates ----------
//
// Inner iteration in FP32 on an FP32 operator: either the FP32 copy of A,
// or (Jacobi and polynomial) the scaled operator A_hat = D^-1/2 A D^-1/2
// with an implicit unit diagonal (off-diagonal values (float)((a_ij s_i) s_j),
// s_i = 1 / sqrt(a_ii) in FP64): Jacobi-PCG as a symmetric scaling, so no
// preconditioner vector is read in the inner loop.  Every dot is
// accumulated in FP64 from exact float x float products.  When the inner
// residual has dropped by `delta` since the last reliable update (or
// reached the tolerance), the inner loop pauses (done = 2): the host then
// adds s o y32 into the FP64 solution, recomputes the true residual
// k - A x in FP64, restarts r32 = (float)(s o r) and keeps p32.

struct MixedScalars {
  double rho;       // <r32, z32>
  double r2;        // <r32, r32>
  double residual;  // sqrt(r2) / inner0, the inner relative residual
  double r2_ref;    // r2 right after the last reliable update
  int iteration;    // inner iterations done
  int done;         // 0 run, 1 stop, 2 reliable update requested
};

struct MixedState {
  MixedScalars cur, next;
  float alpha32;     // (float)alpha of the running iteration: written by the
                     // r update, read by the p update (which also does y)
  double nrm2_0;     // ||k - A x0|| (FP64)
  double inner0;     // ||r32|| at the first reliable update
  double true_res;   // ||k - A x|| / nrm2_0 at the last reliable update
  double tol, delta2;
  int max_iter, updates;
};

// Polynomial preconditioner z = p(A_hat) r (the matrix p(D^-1 A) D^-1 in
// scaled form), p(l) = sum_j c[j] T_j(2 l / b - 1) (Chebyshev basis on
// [0, b]), applied by Clenshaw's recurrence: B_k = c_k r,
// B_j = c_j r + 2 s(A_hat) B_{j+1} - B_{j+2} for j = k-1 .. 1,
// z = c_0 r + s(A_hat) B_1 - B_2 with s(A_hat) v = (2/b) A_hat v - v.
// One SpMV per step.
constexpr int kPolyMaxDegree = 16;
// A reliable update whose true residual fell below this fraction of the
// previous one restarts the search direction (p = z).
constexpr double kResetRatio = 1e-3;
// FP32 ELL with packed slots: slot (k, i) at k * n + i holds the column
// index in its low 32 bits and the value's bits in its high 32 bits, so a
// slot is one 8-byte load and the kernels keep one address per slot instead
// of two (fewer registers: the Clenshaw and SpMV kernels stay at 8 blocks of
// 256 per CU).
struct EllPacked {
  const int *length;
  const uint64_t *slot;
  int n, width;
  int unit;  // 1: implicit unit diagonal, y_i = x_i + (sum of the stored products)
};

// 32-bit byte offsets address every slot of A when its data array has at
// most 2^32 bytes.
template <typename V>
inline bool ell_o32(const Ell<V> &A) {
  return (uint64_t)A.width * (uint64_t)A.n * sizeof(V) <= (1ull << 32);
}

struct PolyParams {
  float c[kPolyMaxDegree + 1];
  float two_over_b;
  int degree;
};

// FP32 diagonal (DIA) layout of an operator whose off-diagonal entries lie
// on at most kDiaMax diagonals j - i = off[d] (off ascending): val[d * n + i]
// holds entry (i, i + off[d]), or the bits kF32Absent where row i has no
// such entry (absent entries are skipped, not multiplied by zero).  No
// column indices and no mask: 4 bytes per slot, and x is read at
// i + off[d], contiguous across a wave.  Slots are addressed by 32-bit byte
// offsets, so dia_fits requires ndiag * n * 4 <= 2^32.  val16 holds the
// same values rounded to FP16 (round to nearest even), row by row: row i's
// record is half_words(ndiag) 32-bit words at val16 + 2 i half_words, value
// d in half d, and an absent entry (or the padding half) holds kHalfAbsent,
// so the polynomial preconditioner's Clenshaw steps, which read them, load
// one record per row instead of one stream per diagonal (2 bytes per slot);
// the CG operator itself stays FP32.
// FP16 bits of a signalling NaN: no FP32 -> FP16 conversion produces it (a
// converted NaN is quiet), so it marks an absent entry of val16.
constexpr uint16_t kHalfAbsent = 0x7D01;
// FP32 bits of a signalling NaN, likewise never produced by an FP64 -> FP32
// conversion: an absent entry of val.  Against a presence-mask byte per row
// (one more stream): 256^3 CG SpMV + <p,q> 111.5 -> 106 us, Clenshaw step on
// FP32 values 110.6 -> 106.1 us.
constexpr uint32_t kF32Absent = 0x7F800001u;
// 32-bit words of one row's FP16 record.
__host__ __device__ inline int half_words(int ndiag) { return (ndiag + 1) / 2; }
struct Dia {
  const float *val;
  const uint16_t *val16;
  int n, ndiag;
  int off[kDiaMax];
  int unit;  // 1: implicit unit diagonal
  // Strip walk of the Clenshaw steps without a dot (dia_strips): the
  // largest offset P is a plane of plane_chunks 256-row chunks, cut into
  // strips of strip_chunks; 0 = not a plane-aligned stencil.
  int plane_chunks, strip_chunks;
  // Symmetric FP16 storage (dia_upper16), or null: when the kSymPairs * 2
  // diagonals come in mirrored pairs (off[d] == -off[ndiag - 1 - d]) and every FP16 value
  // equals its transpose's bit for bit (absent marks included), up16 holds
  // only the upper diagonals, one stream of n halves each: stream k at
  // up16 + k (n + kSymPad) is diagonal ndiag / 2 + k.  Row i reads its lower value of
  // diagonal ndiag / 2 - 1 - k as entry (i - o, i) of stream k (o = off[ndiag
  // / 2 + k]), so the Clenshaw steps load 2 bytes per stored pair instead of
  // 4 and form the same products in the same order.
  const uint16_t *up16;
};
// Diagonal pairs of the symmetric FP16 storage the Clenshaw steps read:
// those of a 3D 7-point stencil.  A compile-time count keeps the row's
// loads free of branches (with the count known only at run time each
// stream's loads sat in a branch of their own and waited for the previous
// stream's: 256^3 step 117 against 75.5 us from the records).
constexpr int kSymPairs = 3;
// Stream k of the symmetric storage starts at up16 + k (n + kSymPad): an
// odd number of 128-byte lines between streams, so that the streams' loads
// for one row do not all fall on one HBM channel when 2n bytes is a large
// power of two.
constexpr int kSymPad = 1088;
// Non-temporal r and B_{j+2} in the Clenshaw steps on the symmetric storage
// once a step's 22 bytes per row outgrow the 256 MB last-level cache
// (per step, degree 12: 224^3 37.8 us cached against 46.7 non-temporal,
// 240^3 61.2 against 58.1, 256^3 75.6 against 67.8; without them the
// 256^3 step read 4 bytes per row less than from the records, PMC
// FETCH_SIZE, in the same time).
inline bool poly_stream_sym(int64_t n) { return 22 * n > (int64_t(1) << 28); }
// Diagonals in mirrored pairs around the (implicit) main diagonal.
inline bool dia_mirrored(const Dia &d) {
  if (d.ndiag < 2 || d.ndiag % 2 != 0 || d.off[d.ndiag / 2] <= 0) return false;
  for (int k = 0; k < d.ndiag; k++)
    if (d.off[k] != -d.off[d.ndiag - 1 - k]) return false;
  return true;
}
// Strip walk geometry of a DIA operator whose largest offset P is a whole
// number of chunks dividing n (a 3D stencil whose x-y plane is P rows):
// strips of the largest divisor of P / 256 not above 32 chunks.  Sets both
// to 0 when the operator does not qualify.
inline void dia_strips(Dia &d) {
  d.plane_chunks = d.strip_chunks = 0;
  if (d.ndiag < 1) return;
  const int64_t P = d.off[d.ndiag - 1];
  if (P <= 0 || P % kBlock != 0 || d.n % P != 0 || d.n / P < 2) return;
  const int pc = (int)(P / kBlock);
  if (pc < 16) return;
  int w = 32;
  while (pc % w) w--;
  d.plane_chunks = pc;
  d.strip_chunks = w;
}
inline bool dia_fits(int64_t ndiag, int64_t n) {
  return ndiag >= 1 && ndiag <= kDiaMax && ndiag * n * 4 <= (1ll << 32) && n * 4 <= (1ll << 32);
}
// DIA is used when its bytes per SpMV (4 ndiag n) are at most those of the
// packed ELL (8 nnz + 4 n).
inline bool dia_pays(int64_t ndiag, int64_t n, int64_t nnz) {
  return dia_fits(ndiag, n) && 4 * ndiag * n <= 8 * nnz + 4 * n;
}

// The FP32 operator of the inner iteration: CRS (row blocks) or, for the
// scaled operator, DIA or ELL (256-row chunks) when they pay.
enum Op32Kind { kOpCrs = 0, kOpEll = 1, kOpDia = 2 };
struct Op32 {
  Crs<float> crs;
  EllPacked ell;
  Dia dia;
  int kind;
};
// Blocks of the SpMV kernels on op (and partials of their dots): DIA with
// up to 4096 blocks, ELL up to 2048, CRS one per row block up to 2048.
inline int op32_grid(const Op32 &op) {
  if (op.kind == kOpDia) return dia_grid(op.dia.n);
  return op.kind == kOpEll ? ell32_grid(op.ell.n) : spmv_grid(op.crs.nrb);
}
// The ELL layout is used when W * n <= 2 (nnz + n), W the longest row.
inline bool ell_pays(int64_t width, int64_t n, int64_t nnz) { return width * n <= 2 * (nnz + n); }

// y = (float)x
hipError_t convert_f64_f32(int64_t n, const double *x, float *y, hipStream_t s);
// y[i] = v
hipError_t fill32(int64_t n, float v, float *y, hipStream_t s);
// Scaled operator set-up.  scaled_diag: s[i] = 1 / sqrt(a_ii) and
// bad[0] = 1 if a row has no or more than one diagonal entry or a_ii <= 0.
// scaled_copy: off-diagonal entries of row i to offset ptr[i] - i.
hipError_t scaled_diag(const Crs<double> &A, double *s, int *bad, hipStream_t st);
hipError_t scaled_copy(const Crs<double> &A, const double *s, int *index_nd, float *value_nd,
                       hipStream_t st);
// The same entries in column-major packed ELL (EllPacked): slot (k, i) at
// k * n + i, length[i].
hipError_t scaled_ell(const Crs<double> &A, const double *s, int *length, uint64_t *slot,
                      hipStream_t st);
// The same entries in DIA (Dia): val and val16 written whole (absent
// entries as kF32Absent and kHalfAbsent); overflow[0] = 1 (zeroed by the
// caller) if a value's magnitude is 65520 or more (its FP16 copy would be
// infinite).
hipError_t scaled_dia(const Crs<double> &A, const double *s, const Dia &D, float *val,
                      uint16_t *val16, int *overflow, hipStream_t st);
// up16 (kSymPairs streams of n halves, kSymPad apart, Dia::up16) from D.val16 of a
// mirrored D; mismatch[0] = 1 (zeroed by the caller) if any FP16 value
// differs from its transpose's or a lower value lies outside the matrix:
// up16 must then not be used.
hipError_t dia_upper16(const Dia &D, uint16_t *up16, int *mismatch, hipStream_t st);
// x64 += scale o (double)y32 (scale nullable: plain add)
hipError_t mp_accumulate(int n, double *x64, const float *y32, const double *scale, hipStream_t s);
// r64 = k - q64; partials.a = ||r64||^2 (pairs order)
hipError_t mp_residual(int n, const double *k, const double *q64, double *r64, Partials pa,
                       hipStream_t s);
// r32 = (float)(scale o r64); y32 = 0; p32 = r32 if first; partials of
// <r32,r32> into pb.a and pb.b (quads order).
hipError_t mp_restart(int n, const double *r64, const double *scale, float *r32, float *y32,
                      float *p32, bool first, Partials pb, hipStream_t s);
// One block: folds ||r64||^2 (pairs, prt.a), <r,r> (quads, prr.a) and <r,z>
// (rz, gz partials) into the restarted state.
hipError_t mp_restart_fold(int n, Partials prt, Partials prr, const double *rz, int gz, bool first,
                           MixedState *st, History h, hipStream_t s);
// Polynomial variant (its r update and restart are mp_update and mp_restart).
// Work buffers w[0..2] hold the Clenshaw B_j (B_j in w[j % 3]; B_k = c_k r is
// recomputed where it is read, not stored); z ends in w[0].
// half: with the DIA layout, the Clenshaw steps read the FP16 values.
hipError_t mp_poly_apply(const Op32 &A, const float *r32, const PolyParams *poly, int degree,
                         float *const w[3], float *pcopy, Partials prz, const MixedState *st,
                         hipStream_t s, bool half);
hipError_t mp_xpay_poly(int n, float *y32, float *p32, const float *z32, Partials prr, Partials prz,
                        int gz, MixedState *st, History h, hipStream_t s);
hipError_t mp_matvec_dot(const Op32 &A, const float *p, float *q, Partials pq, MixedState *st,
                         hipStream_t s);
// alpha = rho / <p,q>; r32 -= alpha q32; partials of <r,r>
hipError_t mp_update(int n, int gpq, float *r32, const float *q32, Partials pq, Partials pr,
                     MixedState *st, History h, hipStream_t s);
// y32 += alpha p32 (the old p: each p element is read once per iteration);
// beta = <r,r>_new / <r,r>; p32 = r32 + beta p32; stop / reliable-update test
hipError_t mp_xpay(int n, float *y32, float *p32, const float *r32, Partials pr, MixedState *st,
                   History h, hipStream_t s);

}  // namespace cghip
