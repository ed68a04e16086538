// SELL-64 storage and kernels for scalar sparse operators.
//
// Rows are grouped in slices of 64 (one wavefront).  Inside a slice the k-th
// entry of every row is stored contiguously (column-major), so each lane
// streams its own row while the wavefront's loads stay fully coalesced: 64
// consecutive columns (256 B) and values (512 B in FP64) per step.  One thread
// owns one row and sums it in the stored (ColumnStencil) order, so with
// EXACT = true the row sums reproduce the reference's sequential loops bit for
// bit (SparseBlockMatrix::vmult, src/LinAlg/sparseblockmatrix.xx:378-392;
// MgInterpolatorMatrix::restrict_zero / prolongate_add,
// src/Discretization/Q1/mginterpolatormatrix.cc:123-147).  Reduced-precision
// (non-exact) sweeps over levels of few slices may instead split a slice over
// K wavefronts (k_sell_split).
#pragma once

#include <type_traits>

#include "gh_kernels.h"

namespace gh {

// ------------------------------------------------------------------ ops
// init(row) gives the accumulator start, step folds one entry, finish writes.
template <typename T, bool EXACT>
__device__ __forceinline__ T fma_step(T acc, T a, T x) {
  if constexpr (EXACT) return add_rn(acc, mul_rn(a, x));
  else return acc + a * x;
}
template <typename T, bool EXACT>
__device__ __forceinline__ T fms_step(T acc, T a, T x) {
  if constexpr (EXACT) return add_rn(acc, -mul_rn(a, x));
  else return acc - a * x;
}

template <typename Y, typename T, bool EXACT>
struct SpmvOp {  // y = d * (A x) + b * y
  static constexpr bool kExact = EXACT;
  Y* y;
  T d, b;
  bool read_y;
  __device__ __forceinline__ T init(int64_t) const { return (T)0; }
  __device__ __forceinline__ T step(T acc, T a, T x) const { return fma_step<T, EXACT>(acc, a, x); }
  __device__ __forceinline__ void finish(int64_t r, T acc) const {
    T v = EXACT ? mul_rn(d, acc) : d * acc;
    if (read_y) v = EXACT ? add_rn(v, mul_rn(b, (T)y[r])) : v + b * (T)y[r];
    y[r] = (Y)v;
  }
};

template <typename T, bool EXACT>
struct ResidualOp {  // r = b - A x: r starts at b and subtracts each a_ij x_j in order
  static constexpr bool kExact = EXACT;
  T* r;
  const T* b;
  __device__ __forceinline__ T init(int64_t i) const { return b[i]; }
  __device__ __forceinline__ T step(T acc, T a, T x) const { return fms_step<T, EXACT>(acc, a, x); }
  __device__ __forceinline__ void finish(int64_t i, T acc) const { r[i] = acc; }
};

template <typename T, typename DT, bool EXACT>
struct JacobiOp {  // xn = x + omega * (dinv * (b - A x))
  static constexpr bool kExact = EXACT;
  T* xn;
  const T* x;
  const T* b;
  const DT* dinv;
  T omega;
  int richardson;
  __device__ __forceinline__ T init(int64_t i) const { return b[i]; }
  __device__ __forceinline__ T step(T acc, T a, T xx) const { return fms_step<T, EXACT>(acc, a, xx); }
  __device__ __forceinline__ void finish(int64_t i, T h) const {
    if (!richardson) h = EXACT ? mul_rn((T)dinv[i], h) : (T)dinv[i] * h;
    xn[i] = EXACT ? add_rn(x[i], mul_rn(omega, h)) : x[i] + omega * h;
  }
};

template <typename T>
struct RestrictOp {  // uc = R uf (packed weights)
  T* uc;
  __device__ __forceinline__ T init(int64_t) const { return (T)0; }
  __device__ __forceinline__ T step(T acc, T w, T x) const { return add_rn(acc, mul_rn(w, x)); }
  __device__ __forceinline__ void finish(int64_t i, T acc) const { uc[i] = acc; }
};

// uc = R uf, and the coarse level's first smoothing sweep from a zero start
// in the same pass: ux = 0 + omega * (dinv * (uc - 0)), the operations of a
// Jacobi (or Richardson) sweep whose iterate is zero
template <typename T, typename DT>
struct RestrictSmoothOp {
  T* uc;
  T* ux;
  const DT* dinv;
  T omega;
  int richardson;
  __device__ __forceinline__ T init(int64_t) const { return (T)0; }
  __device__ __forceinline__ T step(T acc, T w, T x) const { return add_rn(acc, mul_rn(w, x)); }
  __device__ __forceinline__ void finish(int64_t i, T acc) const {
    uc[i] = acc;
    T h = acc - (T)0;
    if (!richardson) h = (T)dinv[i] * h;
    ux[i] = (T)0 + omega * h;
  }
};

template <typename T>
struct ProlongAddOp {  // uf += P uc, term by term
  T* uf;
  __device__ __forceinline__ T init(int64_t i) const { return uf[i]; }
  __device__ __forceinline__ T step(T acc, T w, T x) const { return add_rn(acc, mul_rn(w, x)); }
  __device__ __forceinline__ void finish(int64_t i, T acc) const { uf[i] = acc; }
};

template <typename T, typename UT>
struct CorrectOp {  // v = P' uc ; u += mgomega * v  (empty rows leave u as it is)
  UT* u;
  T mgomega;
  __device__ __forceinline__ T init(int64_t) const { return (T)0; }
  __device__ __forceinline__ T step(T acc, T w, T x) const { return add_rn(acc, mul_rn(w, x)); }
  __device__ __forceinline__ void finish(int64_t i, T v) const {
    u[i] = (UT)add_rn((T)u[i], mul_rn(mgomega, v));
  }
};

// ---------------------------------------------------------------- kernels
// Workgroups are dealt to the 8 XCDs round-robin (workgroup i runs on XCD
// i mod 8).  Remapping the workgroup index gives every XCD one contiguous
// range of rows, so the x entries a row range gathers are fetched into that
// XCD's own L2 instead of into all eight.
constexpr unsigned kXcds = 8;
__device__ __forceinline__ unsigned xcd_block(unsigned b, unsigned G) {
  const unsigned q = G / kXcds, r = G % kXcds;
  const unsigned xcd = b % kXcds, k = b / kXcds;
  return xcd * q + (xcd < r ? xcd : r) + k;
}

// ------------------------------------------------- block matrices (BSell)
// B consecutive values of one x block, widened to the accumulator type;
// 16-byte vector loads where the data allow it (the host checks the alignment
// before it picks this kernel).
template <int B, typename T, typename Acc>
__device__ __forceinline__ void load_block_row(const T* __restrict__ p, Acc (&o)[B]) {
  if constexpr (B == 4 && std::is_same<T, double>::value) {
    typedef double d2 __attribute__((ext_vector_type(2)));
    const d2 a = *reinterpret_cast<const d2*>(p), b = *reinterpret_cast<const d2*>(p + 2);
    o[0] = (Acc)a.x; o[1] = (Acc)a.y; o[2] = (Acc)b.x; o[3] = (Acc)b.y;
  } else if constexpr (B == 4 && std::is_same<T, float>::value) {
    typedef float f4 __attribute__((ext_vector_type(4)));
    const f4 a = *reinterpret_cast<const f4*>(p);
    o[0] = (Acc)a.x; o[1] = (Acc)a.y; o[2] = (Acc)a.z; o[3] = (Acc)a.w;
  } else {
#pragma unroll
    for (int c = 0; c < B; ++c) o[c] = (Acc)ld(p + c);
  }
}

// One lane per component row t = i*B + r: for each kept block of row i (in
// the row's column order) sum = sum_c A(r, c) x_c from zero, then
// acc = op.step(acc, sum) -- the per-block sums of k_bsr, in the same order,
// so both kernels return the same bits.  Only all-zero block rows are left
// out (their sum is an exact zero).
template <int B, typename VT, typename XT, typename Acc, class Op>
__global__ __launch_bounds__(kBlock) void k_bsell(int64_t nslices, const uint32_t* __restrict__ sptr,
                                                  const int32_t* __restrict__ row,
                                                  const uint16_t* __restrict__ rlen,
                                                  const int32_t* __restrict__ col,
                                                  const VT* __restrict__ val,
                                                  const XT* __restrict__ x, Op op) {
  const int64_t g = (int64_t)xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
  const int64_t s = g >> 6;
  if (s >= nslices) return;
  const int32_t t = row[g];
  if (t < 0) return;
  const int len = rlen[g];
  const int64_t p0 = (int64_t)sptr[s] * 64 + (threadIdx.x & 63);
  Acc acc = op.init(t);
  const int lane = threadIdx.x & 63;
  const int64_t v0 = (int64_t)sptr[s] * B * 64 + lane;  // value c of entry k: v0 + (k*B + c)*64
  int k = 0;
  // Four entries per step.  The step's value and x loads are all issued
  // before its arithmetic (the scheduling barrier keeps the compiler from
  // serialising the x gathers to save registers), and the next step's columns
  // are fetched one step ahead.
  int32_t c[4];
  if (len >= 4) {
#pragma unroll
    for (int u = 0; u < 4; ++u) c[u] = col[p0 + (int64_t)u * 64];
  }
  for (; k + 4 <= len; k += 4) {
    Acc m[4][B], q[4][B];
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int cc = 0; cc < B; ++cc) m[u][cc] = (Acc)ld(val + v0 + ((int64_t)(k + u) * B + cc) * 64);
#pragma unroll
    for (int u = 0; u < 4; ++u) load_block_row<B>(x + (int64_t)c[u] * B, q[u]);
    if (k + 8 <= len) {
#pragma unroll
      for (int u = 0; u < 4; ++u) c[u] = col[p0 + (int64_t)(k + 4 + u) * 64];
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      Acc sum = (Acc)0;
#pragma unroll
      for (int cc = 0; cc < B; ++cc) sum = add_rn(sum, mul_rn(m[u][cc], q[u][cc]));
      acc = op.step(acc, sum);
    }
  }
  for (; k < len; ++k) {
    Acc m[B], q[B];
#pragma unroll
    for (int cc = 0; cc < B; ++cc) m[cc] = (Acc)ld(val + v0 + ((int64_t)k * B + cc) * 64);
    load_block_row<B>(x + (int64_t)col[p0 + (int64_t)k * 64] * B, q);
    Acc sum = (Acc)0;
#pragma unroll
    for (int cc = 0; cc < B; ++cc) sum = add_rn(sum, mul_rn(m[cc], q[cc]));
    acc = op.step(acc, sum);
  }
  op.finish(t, acc);
}

// One row of a valued operator: entries k = 0..len-1 at stride kSlice, column
// = cb + cp[k * kSlice] (16-bit offsets) or cp[k * kSlice] (32-bit, cb = 0).
// With K > 1 the caller folds every K-th entry only: len counts those, and
// cp and vp point at the first of them.  The loads of 8 (then 4) entries are
// issued before their adds, which stay in entry order: a long row pays one
// column -> x round trip per 8 entries, and the sum keeps its bits.
template <int K = 1, typename CT, typename VT, typename XT, typename T, class Op>
__device__ __forceinline__ T sell_row(const CT* __restrict__ cp, uint32_t cb,
                                      const VT* __restrict__ vp, int len,
                                      const XT* __restrict__ x, T acc, const Op& op) {
  constexpr int S = K * kSlice;
  int k = 0;
  for (; k + 8 <= len; k += 8) {
    uint32_t c[8];
    T a[8], xv[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) c[u] = cb + cp[(k + u) * S];
#pragma unroll
    for (int u = 0; u < 8; ++u) a[u] = (T)ld(vp + (k + u) * S);
#pragma unroll
    for (int u = 0; u < 8; ++u) xv[u] = (T)x[c[u]];
#pragma unroll
    for (int u = 0; u < 8; ++u) acc = op.step(acc, a[u], xv[u]);
  }
  for (; k + 4 <= len; k += 4) {
    const uint32_t c0 = cb + cp[(k + 0) * S], c1 = cb + cp[(k + 1) * S];
    const uint32_t c2 = cb + cp[(k + 2) * S], c3 = cb + cp[(k + 3) * S];
    const T a0 = (T)ld(vp + (k + 0) * S), a1 = (T)ld(vp + (k + 1) * S);
    const T a2 = (T)ld(vp + (k + 2) * S), a3 = (T)ld(vp + (k + 3) * S);
    const T x0 = (T)x[c0], x1 = (T)x[c1], x2 = (T)x[c2], x3 = (T)x[c3];
    acc = op.step(acc, a0, x0);
    acc = op.step(acc, a1, x1);
    acc = op.step(acc, a2, x2);
    acc = op.step(acc, a3, x3);
  }
  for (; k < len; ++k) acc = op.step(acc, (T)ld(vp + k * S), (T)x[cb + cp[k * S]]);
  return acc;
}

// One row that follows a shared relative pattern: column = row + off[k]
// (the offsets come from a table of a few KB, served by the caches).  With
// K > 1, as above: every K-th entry, off and vp at the first of them.
template <int K = 1, typename VT, typename XT, typename T, class Op>
__device__ __forceinline__ T sell_row_pattern(const int32_t* __restrict__ off, uint32_t row,
                                              const VT* __restrict__ vp, int len,
                                              const XT* __restrict__ x, T acc, const Op& op) {
  constexpr int S = K * kSlice;
  int k = 0;
  for (; k + 8 <= len; k += 8) {
    uint32_t c[8];
    T a[8], xv[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) c[u] = row + off[(k + u) * K];
#pragma unroll
    for (int u = 0; u < 8; ++u) a[u] = (T)ld(vp + (k + u) * S);
#pragma unroll
    for (int u = 0; u < 8; ++u) xv[u] = (T)x[c[u]];
#pragma unroll
    for (int u = 0; u < 8; ++u) acc = op.step(acc, a[u], xv[u]);
  }
  for (; k + 4 <= len; k += 4) {
    const uint32_t c0 = row + off[(k + 0) * K], c1 = row + off[(k + 1) * K];
    const uint32_t c2 = row + off[(k + 2) * K], c3 = row + off[(k + 3) * K];
    const T a0 = (T)ld(vp + (k + 0) * S), a1 = (T)ld(vp + (k + 1) * S);
    const T a2 = (T)ld(vp + (k + 2) * S), a3 = (T)ld(vp + (k + 3) * S);
    const T x0 = (T)x[c0], x1 = (T)x[c1], x2 = (T)x[c2], x3 = (T)x[c3];
    acc = op.step(acc, a0, x0);
    acc = op.step(acc, a1, x1);
    acc = op.step(acc, a2, x2);
    acc = op.step(acc, a3, x3);
  }
  for (; k < len; ++k) acc = op.step(acc, (T)ld(vp + k * S), (T)x[row + off[k * K]]);
  return acc;
}

// valued operator (matrix); VT storage type, T arithmetic type.  A slice is one
// wavefront, so the column-encoding branch is uniform across the wave.
template <typename VT, typename XT, typename T, class Op>
__global__ __launch_bounds__(kBlock) void k_sell(int64_t n, const uint32_t* __restrict__ sptr,
                                                 const uint16_t* __restrict__ rlen,
                                                 const uint16_t* __restrict__ cidx,
                                                 const uint32_t* __restrict__ cptr,
                                                 const uint32_t* __restrict__ cbase,
                                                 const int32_t* __restrict__ ptab,
                                                 const int32_t* __restrict__ pstart,
                                                 const VT* __restrict__ val,
                                                 const int32_t* __restrict__ perm,
                                                 const XT* __restrict__ x, Op op) {
  // Lanes past the last slot repeat slot n-1 with their own (padding) slots
  // and only skip the store, so the control flow stays uniform across the wave
  // and wave-uniform loads can use the scalar unit.  Slot t runs row perm[t]
  // (t itself without a permutation).
  const int64_t slot0 = (int64_t)xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
  const bool active = slot0 < n;
  const int64_t slot = active ? slot0 : n - 1;
  const int64_t row = perm ? (int64_t)perm[slot] : slot;
  const int lane = threadIdx.x & (kSlice - 1);
  // the slice index is the same for the 64 lanes of a wavefront
  const int64_t s = __builtin_amdgcn_readfirstlane((int)(slot >> 6));
  const VT* vp = val + sptr[s] + lane;
  const uint32_t meta = cptr[s];
  T acc = op.init(row);
  if ((meta & 0x60000000u) == 0x60000000u) {
    // one pattern for the whole slice: wave-uniform offsets (scalar loads)
    const int pid = __builtin_amdgcn_readfirstlane((int)cbase[s]);
    const int p0 = pstart[pid];
    acc = sell_row_pattern(ptab + p0, (uint32_t)row, vp, pstart[pid + 1] - p0, x, acc, op);
  } else {
    const int r = rlen[slot];  // entry count, or pattern id in pattern slices
    if (meta & 0x40000000u) {
      const int p0 = pstart[r];
      acc = sell_row_pattern(ptab + p0, (uint32_t)row, vp, pstart[r + 1] - p0, x, acc, op);
    } else if (meta & 0x80000000u) {
      acc = sell_row((const uint32_t*)(cidx + (meta & 0x1FFFFFFFu)) + lane, 0u, vp, r, x, acc, op);
    } else {
      acc = sell_row(cidx + meta + lane, cbase[s], vp, r, x, acc, op);
    }
  }
  if (active) op.finish(row, acc);
}

// The same sweep for reduced-precision (non-exact) ops on levels with few
// slices, where one thread per row leaves the chip mostly idle and each row
// is a long chain of dependent loads (Q2 rows hold up to 125 entries).  A
// workgroup of K wavefronts takes one slice: wavefront w folds entries
// w, w + K, w + 2K, ... of each of the slice's 64 rows (every load still
// covers 64 contiguous slots), and wavefront 0 adds the K partial sums, in
// order, to the row's start value.  Each partial chain starts from zero,
// which is right for ops whose step is acc + a x or acc - a x (SpmvOp,
// ResidualOp, JacobiOp).  That reorders the row sum, so exact ops are
// rejected at compile time.  Every thread reaches the barrier: lanes past
// the last row repeat row n-1, as in k_sell, and only skip the store.
template <typename VT, typename XT, typename T, class Op, int K>
__global__ __launch_bounds__(kSlice * K) void k_sell_split(
    int64_t n, const uint32_t* __restrict__ sptr, const uint16_t* __restrict__ rlen,
    const uint16_t* __restrict__ cidx, const uint32_t* __restrict__ cptr,
    const uint32_t* __restrict__ cbase, const int32_t* __restrict__ ptab,
    const int32_t* __restrict__ pstart, const VT* __restrict__ val,
    const int32_t* __restrict__ perm, const XT* __restrict__ x, Op op) {
  static_assert(!Op::kExact, "split sweeps reorder the row sums");
  __shared__ T part[K][kSlice];
  const int64_t s = xcd_block(blockIdx.x, gridDim.x);  // one slice per workgroup
  const int lane = threadIdx.x & (kSlice - 1);
  const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / kSlice));
  const int64_t slot0 = s * kSlice + lane;
  const bool active = slot0 < n;
  const int64_t slot = active ? slot0 : n - 1;
  const int64_t row = perm ? (int64_t)perm[slot] : slot;
  const VT* vp = val + sptr[s] + lane + w * kSlice;
  const uint32_t meta = cptr[s];
  // entries w, w + K, ... below len
  auto mine = [w](int len) { return len > w ? (len - w + K - 1) / K : 0; };
  T acc = (T)0;
  if ((meta & 0x60000000u) == 0x60000000u) {
    const int pid = __builtin_amdgcn_readfirstlane((int)cbase[s]);
    const int p0 = pstart[pid];
    acc = sell_row_pattern<K>(ptab + p0 + w, (uint32_t)row, vp, mine(pstart[pid + 1] - p0), x,
                              acc, op);
  } else {
    const int r = rlen[slot];
    if (meta & 0x40000000u) {
      const int p0 = pstart[r];
      acc = sell_row_pattern<K>(ptab + p0 + w, (uint32_t)row, vp, mine(pstart[r + 1] - p0), x,
                                acc, op);
    } else if (meta & 0x80000000u) {
      acc = sell_row<K>((const uint32_t*)(cidx + (meta & 0x1FFFFFFFu)) + lane + w * kSlice, 0u,
                        vp, mine(r), x, acc, op);
    } else {
      acc = sell_row<K>(cidx + meta + lane + w * kSlice, cbase[s], vp, mine(r), x, acc, op);
    }
  }
  part[w][lane] = acc;
  __syncthreads();
  if (w == 0) {
    T sum = op.init(row);
#pragma unroll
    for (int i = 0; i < K; ++i) sum += part[i][lane];
    if (active) op.finish(row, sum);
  }
}

// Split factor of a reduced-precision sweep: double K while the level runs
// fewer than 8192 wavefronts (32 per CU) and every wavefront still gets at
// least 8 of the longest row's entries, up to 16 wavefronts per slice.
inline int sell_split(const Sell& S) {
  int K = 1;
  while (K < 16 && S.nslices * K < 8192 && S.maxlen >= 16 * K) K *= 2;
  return K;
}

// packed-weight operator (grid transfer): w = 2^-code from the column's top bits
template <typename T, class Op>
__global__ __launch_bounds__(kBlock) void k_sell_w(int64_t n, const uint32_t* __restrict__ sptr,
                                                   const uint16_t* __restrict__ rlen,
                                                   const uint32_t* __restrict__ col,
                                                   const T* __restrict__ x, Op op) {
  const int64_t row = (int64_t)xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
  if (row >= n) return;
  const int lane = threadIdx.x & (kSlice - 1);
  const uint32_t* cp = col + sptr[row >> 6] + lane;
  const int len = rlen[row];
  T acc = op.init(row);
  int k = 0;
  for (; k + 4 <= len; k += 4) {
    const uint32_t e0 = cp[(k + 0) * kSlice], e1 = cp[(k + 1) * kSlice];
    const uint32_t e2 = cp[(k + 2) * kSlice], e3 = cp[(k + 3) * kSlice];
    const T x0 = x[e0 & kIdxMask], x1 = x[e1 & kIdxMask], x2 = x[e2 & kIdxMask],
            x3 = x[e3 & kIdxMask];
    acc = op.step(acc, code_weight<T>(e0 >> 29), x0);
    acc = op.step(acc, code_weight<T>(e1 >> 29), x1);
    acc = op.step(acc, code_weight<T>(e2 >> 29), x2);
    acc = op.step(acc, code_weight<T>(e3 >> 29), x3);
  }
  for (; k < len; ++k) {
    const uint32_t e = cp[k * kSlice];
    acc = op.step(acc, code_weight<T>(e >> 29), x[e & kIdxMask]);
  }
  op.finish(row, acc);
}

// The grid transfers of an FP32 V-cycle on levels of few slices, split as
// k_sell_split splits the sweeps: wavefront w of a slice's K folds entries
// w, w + K, ... of each row from zero, and wavefront 0 adds the K partial sums,
// in order, to op.init(row).  Every transfer op steps acc + w x, so only the
// order of the row sum changes; the FP64 transfers keep one thread per row and
// their bits.
template <typename T, class Op, int K>
__global__ __launch_bounds__(kSlice * K) void k_sell_w_split(int64_t n,
                                                             const uint32_t* __restrict__ sptr,
                                                             const uint16_t* __restrict__ rlen,
                                                             const uint32_t* __restrict__ col,
                                                             const T* __restrict__ x, Op op) {
  __shared__ T part[K][kSlice];
  const int64_t s = xcd_block(blockIdx.x, gridDim.x);  // one slice per workgroup
  const int lane = threadIdx.x & (kSlice - 1);
  const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / kSlice));
  const int64_t row0 = s * kSlice + lane;
  const bool active = row0 < n;
  const int64_t row = active ? row0 : n - 1;
  const uint32_t* cp = col + sptr[s] + lane + w * kSlice;
  const int len = rlen[row];
  const int mine = len > w ? (len - w + K - 1) / K : 0;
  constexpr int S = K * kSlice;
  T acc = (T)0;
  int k = 0;
  for (; k + 4 <= mine; k += 4) {
    const uint32_t e0 = cp[(k + 0) * S], e1 = cp[(k + 1) * S];
    const uint32_t e2 = cp[(k + 2) * S], e3 = cp[(k + 3) * S];
    const T x0 = x[e0 & kIdxMask], x1 = x[e1 & kIdxMask], x2 = x[e2 & kIdxMask],
            x3 = x[e3 & kIdxMask];
    acc = op.step(acc, code_weight<T>(e0 >> 29), x0);
    acc = op.step(acc, code_weight<T>(e1 >> 29), x1);
    acc = op.step(acc, code_weight<T>(e2 >> 29), x2);
    acc = op.step(acc, code_weight<T>(e3 >> 29), x3);
  }
  for (; k < mine; ++k) {
    const uint32_t e = cp[k * S];
    acc = op.step(acc, code_weight<T>(e >> 29), x[e & kIdxMask]);
  }
  part[w][lane] = acc;
  __syncthreads();
  if (w == 0) {
    T sum = op.init(row);
#pragma unroll
    for (int i = 0; i < K; ++i) sum += part[i][lane];
    if (active) op.finish(row, sum);
  }
}

// Split factor of an FP32 transfer: double K while the doubled level still
// runs at most 8192 wavefronts and every wavefront still gets at least 4 of
// the longest row's entries, up to 16.  (Levels of more than 4096 slices stay
// unsplit: at 64^3 a split of 2 measured slower than one thread per row.)
inline int sell_w_split(const Sell& S) {
  int K = 1;
  while (K < 16 && S.nslices * 2 * K <= 8192 && S.maxlen >= 4 * K) K *= 2;
  return K;
}

template <typename T, class Op, int K>
inline int sell_w_split_launch(gh_ctx* ctx, const Sell& S, const T* x, Op op) {
  hipLaunchKernelGGL((k_sell_w_split<T, Op, K>), dim3((unsigned)S.nslices), dim3(kSlice * K), 0,
                     ctx->stream, S.n, S.sptr, S.rlen, S.col, x, op);
  GH_HIP(hipGetLastError());
  return GH_OK;
}

template <typename VT, typename XT, typename T, class Op, int K>
inline int sell_split_launch(gh_ctx* ctx, const Sell& S, const XT* x, Op op) {
  hipLaunchKernelGGL((k_sell_split<VT, XT, T, Op, K>), dim3((unsigned)S.nslices),
                     dim3(kSlice * K), 0, ctx->stream, S.n, S.sptr, S.rlen, S.cidx, S.cptr,
                     S.cbase, S.ptab, S.pstart, (const VT*)S.val, S.perm, x, op);
  GH_HIP(hipGetLastError());
  return GH_OK;
}

template <typename VT, typename XT, typename T, class Op>
inline int sell_launch(gh_ctx* ctx, const Sell& S, const XT* x, Op op) {
  if (S.n == 0) return GH_OK;
  if constexpr (!Op::kExact) {
    switch (sell_split(S)) {
      case 2: return sell_split_launch<VT, XT, T, Op, 2>(ctx, S, x, op);
      case 4: return sell_split_launch<VT, XT, T, Op, 4>(ctx, S, x, op);
      case 8: return sell_split_launch<VT, XT, T, Op, 8>(ctx, S, x, op);
      case 16: return sell_split_launch<VT, XT, T, Op, 16>(ctx, S, x, op);
      default: break;
    }
  }
  hipLaunchKernelGGL((k_sell<VT, XT, T, Op>), dim3(grid_for(S.n)), dim3(kBlock), 0, ctx->stream,
                     S.n, S.sptr, S.rlen, S.cidx, S.cptr, S.cbase, S.ptab, S.pstart,
                     (const VT*)S.val, S.perm, x, op);
  GH_HIP(hipGetLastError());
  return GH_OK;
}

template <typename XT, typename T, class Op>
inline int sell_dispatch(gh_ctx* ctx, const Sell& S, const XT* x, Op op) {
  if (S.prec == GH_F64) return sell_launch<double, XT, T, Op>(ctx, S, x, op);
  if (S.prec == GH_F32) return sell_launch<float, XT, T, Op>(ctx, S, x, op);
  return sell_launch<__half, XT, T, Op>(ctx, S, x, op);
}

template <typename T, class Op>
inline int sell_w_launch(gh_ctx* ctx, const Sell& S, const T* x, Op op) {
  if (S.n == 0) return GH_OK;
  if constexpr (std::is_same<T, float>::value) {
    switch (sell_w_split(S)) {
      case 2: return sell_w_split_launch<T, Op, 2>(ctx, S, x, op);
      case 4: return sell_w_split_launch<T, Op, 4>(ctx, S, x, op);
      case 8: return sell_w_split_launch<T, Op, 8>(ctx, S, x, op);
      case 16: return sell_w_split_launch<T, Op, 16>(ctx, S, x, op);
      default: break;
    }
  }
  hipLaunchKernelGGL((k_sell_w<T, Op>), dim3(grid_for(S.n)), dim3(kBlock), 0, ctx->stream, S.n,
                     S.sptr, S.rlen, S.col, x, op);
  GH_HIP(hipGetLastError());
  return GH_OK;
}

}  // namespace gh
