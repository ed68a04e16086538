// Device kernels of the gascoignelib_amd solve path (gfx950, wave64).
//
// All kernels are bandwidth kernels: no MFMA.  Arithmetic types:
//   Acc = double on the FP64 path, float on the mixed path;
//   matrix values may be stored as double, float or __half.
#pragma once

#include "gh_internal.h"

namespace gh {

template <typename T> __device__ __forceinline__ T to_acc(double v) { return (T)v; }

__device__ __forceinline__ float ld(const __half* p) { return __half2float(*p); }
__device__ __forceinline__ float ld(const float* p) { return *p; }
__device__ __forceinline__ double ld(const double* p) { return *p; }

// Separately rounded multiply and add: the operations carry no contraction
// flag, so they are never fused into an FMA, even after inlining.  Used where
// the reference's sequential CPU expression must be reproduced bit for bit.
template <typename T> __device__ __forceinline__ T mul_rn(T a, T b) {
#pragma clang fp contract(off)
  return a * b;
}
template <typename T> __device__ __forceinline__ T add_rn(T a, T b) {
#pragma clang fp contract(off)
  return a + b;
}

// weight code of packed transfer columns: w = 2^-code (1, 1/2, 1/4, 1/8)
constexpr uint32_t kIdxMask = 0x1FFFFFFFu;
template <typename T> __device__ __forceinline__ T code_weight(uint32_t c) {
  return c == 0 ? (T)1.0 : c == 1 ? (T)0.5 : c == 2 ? (T)0.25 : (T)0.125;
}

// ------------------------------------------------------------- reductions
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_down(v, off, 64));
  return v;
}

// block of kBlock threads; result valid in thread 0
template <bool MAX>
__device__ __forceinline__ double block_reduce(double v) {
  __shared__ double s[kBlock / 64];
  v = MAX ? wave_max(v) : wave_sum(v);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) s[wave] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0) {
    r = s[0];
#pragma unroll
    for (int w = 1; w < kBlock / 64; ++w) r = MAX ? fmax(r, s[w]) : r + s[w];
  }
  __syncthreads();
  return r;
}

// Reductions run in two phases without fences or atomics: every block of the
// producing kernel stores one partial (kRedBlocks of them per slot), and the
// consumer folds the partials in a fixed order (sum_partials).  The grid is
// the same for every call, so results are bitwise reproducible.
template <bool MAX>
__device__ __forceinline__ double sum_partials(const double* __restrict__ part) {
  double a = 0.0;
  for (int i = threadIdx.x; i < kRedBlocks; i += kBlock) a = MAX ? fmax(a, part[i]) : a + part[i];
  return block_reduce<MAX>(a);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_dot_part(int64_t n, const T* __restrict__ x,
                                                     const T* __restrict__ y,
                                                     double* __restrict__ part) {
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
  for (; i + 3 * stride < n; i += 4 * stride) {
    a0 += (double)x[i] * (double)y[i];
    a1 += (double)x[i + stride] * (double)y[i + stride];
    a2 += (double)x[i + 2 * stride] * (double)y[i + 2 * stride];
    a3 += (double)x[i + 3 * stride] * (double)y[i + 3 * stride];
  }
  for (; i < n; i += stride) a0 += (double)x[i] * (double)y[i];
  const double acc = block_reduce<false>((a0 + a1) + (a2 + a3));
  if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_amax_part(int64_t n, const T* __restrict__ x,
                                                      double* __restrict__ part) {
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
  for (; i + 3 * stride < n; i += 4 * stride) {
    a0 = fmax(a0, fabs((double)x[i]));
    a1 = fmax(a1, fabs((double)x[i + stride]));
    a2 = fmax(a2, fabs((double)x[i + 2 * stride]));
    a3 = fmax(a3, fabs((double)x[i + 3 * stride]));
  }
  for (; i < n; i += stride) a0 = fmax(a0, fabs((double)x[i]));
  const double acc = block_reduce<true>(fmax(fmax(a0, a1), fmax(a2, a3)));
  if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

// one block per slot: out[s] = fold of the slot's partials
template <bool MAX>
__global__ __launch_bounds__(kBlock) void k_finalize(const double* __restrict__ part,
                                                     double* __restrict__ out) {
  const double a = sum_partials<MAX>(part + (int64_t)blockIdx.x * kRedBlocks);
  if (threadIdx.x == 0) out[blockIdx.x] = a;
}

// ----------------------------------------------------------- vector kernels
template <typename T>
__global__ __launch_bounds__(kBlock) void k_axpby(int64_t n, T* __restrict__ y, T a, T d,
                                                  const T* __restrict__ x, int mode) {
  // mode 0: y = a*y + d*x ; 1: y += d*x ; 2: y = d*x -- each product rounded
  // before the sum, as nvector's loops compute them (src/Common/nvector.h:
  // equ / add / sadd), not fused into one multiply-add
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock) {
    if (mode == 0)
      y[i] = add_rn(mul_rn(a, y[i]), mul_rn(d, x[i]));
    else if (mode == 1)
      y[i] = add_rn(y[i], mul_rn(d, x[i]));
    else
      y[i] = mul_rn(d, x[i]);
  }
}

template <typename D, typename S>
__global__ __launch_bounds__(kBlock) void k_convert(int64_t n, D* __restrict__ dst,
                                                    const S* __restrict__ src) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock)
    dst[i] = (D)src[i];
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_fill0(int64_t n, T* __restrict__ x) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock)
    x[i] = (T)0;
}

// ------------------------------------------------------ block SpMV (B > 1)
// SparseBlockMatrix<FMatrixBlock<B>>::vmult (src/LinAlg/sparseblockmatrix.xx:
// 378-392) adds, block by block in stored order, s * (row r of the block times
// x_col) into y (FMatrixBlock::cadd, src/LinAlg/fmatrixblock.h); the ops below
// take the same per-block sums in the same order, without contraction.
template <typename Y, typename Acc>
struct BsrSpmvOp {  // y = d * A x + b * y
  Y* y;
  Acc d, b;
  bool read_y;
  __device__ __forceinline__ Acc init(int64_t t) const {
    if (!read_y) return (Acc)0;
    return b == (Acc)1 ? (Acc)y[t] : mul_rn(b, (Acc)y[t]);
  }
  __device__ __forceinline__ Acc step(Acc acc, Acc sum) const {
    return add_rn(acc, mul_rn(d, sum));
  }
  __device__ __forceinline__ void finish(int64_t t, Acc acc) const { y[t] = (Y)acc; }
};

template <typename T>
struct BsrResidualOp {  // r = b - A x (MatrixResidual: r = b; vmult(r, x, -1))
  T* r;
  const T* b;
  __device__ __forceinline__ T init(int64_t t) const { return b[t]; }
  __device__ __forceinline__ T step(T acc, T sum) const { return add_rn(acc, -sum); }
  __device__ __forceinline__ void finish(int64_t t, T acc) const { r[t] = acc; }
};

// One thread per block row component (i, r)
template <typename VT, typename XT, typename Acc, class Op>
__global__ __launch_bounds__(kBlock) void k_bsr(int64_t n, int B, const int32_t* __restrict__ rowptr,
                                                const int32_t* __restrict__ col,
                                                const VT* __restrict__ val,
                                                const XT* __restrict__ x, Op op) {
  const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (t >= n * B) return;
  const int64_t i = t / B;
  const int r = (int)(t - i * B);
  Acc acc = op.init(t);
  for (int pos = rowptr[i]; pos < rowptr[i + 1]; ++pos) {
    const VT* m = val + (int64_t)pos * B * B + (int64_t)r * B;
    const XT* q = x + (int64_t)col[pos] * B;
    Acc sum = (Acc)0;
    for (int c = 0; c < B; ++c) sum = add_rn(sum, mul_rn((Acc)ld(m + c), (Acc)q[c]));
    acc = op.step(acc, sum);
  }
  op.finish(t, acc);
}

// First sweep from x = 0 (no product with A): h = b - 0, h = dinv h,
// xn = 0 + omega h, the operations of a sweep whose iterate is zero
template <typename T, typename DT>
__global__ __launch_bounds__(kBlock) void k_jacobi_zero(int64_t n, T* __restrict__ xn,
                                                        const T* __restrict__ b,
                                                        const DT* __restrict__ dinv, T omega,
                                                        int richardson) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock) {
    T h = b[i] - (T)0;
    if (!richardson) h = (T)dinv[i] * h;
    xn[i] = (T)0 + omega * h;
  }
}

// block Jacobi: xn(i) = x(i) + omega * Dinv_i (b(i) - (A x)(i)); one thread per row
template <typename VT, typename T, typename DT>
__global__ __launch_bounds__(kBlock) void k_bsr_jacobi(int64_t n, int B,
                                                       const int32_t* __restrict__ rowptr,
                                                       const int32_t* __restrict__ col,
                                                       const VT* __restrict__ val,
                                                       const T* __restrict__ x,
                                                       const T* __restrict__ b,
                                                       const DT* __restrict__ dinv,
                                                       T* __restrict__ xn, T omega, int x_zero) {
  // h = b - A x (block by block), h <- Dinv_i h (NodeMatrix product, sums from
  // zero in column order), x + omega h: each step as the CPU forms it
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  T h[8];
  for (int r = 0; r < B; ++r) h[r] = b[i * B + r];
  if (!x_zero) {
    for (int pos = rowptr[i]; pos < rowptr[i + 1]; ++pos) {
      const VT* m = val + (int64_t)pos * B * B;
      const T* q = x + (int64_t)col[pos] * B;
      for (int r = 0; r < B; ++r) {
        T sum = (T)0;
        for (int c = 0; c < B; ++c) sum = add_rn(sum, mul_rn((T)ld(m + r * B + c), q[c]));
        h[r] = add_rn(h[r], -sum);
      }
    }
  }
  const DT* D = dinv + i * B * B;
  for (int r = 0; r < B; ++r) {
    T s = (T)0;
    for (int c = 0; c < B; ++c) s = add_rn(s, mul_rn((T)D[r * B + c], h[c]));
    xn[i * B + r] = add_rn(x_zero ? (T)0 : x[i * B + r], mul_rn(omega, s));
  }
}

// In place: m (B x B, row-major, in registers) becomes its inverse, by the
// operations of NodeMatrix::gauss_jordan (src/LinAlg/nodematrix.h:159-213),
// uncontracted.  Returns false for a zero pivot (m is then undefined).
template <int B>
__device__ __forceinline__ bool gauss_jordan(double (&m)[B * B]) {
#pragma clang fp contract(off)
  int p[B];
#define M_(a, c) m[(a) * B + (c)]
#pragma unroll
  for (int k = 0; k < B; ++k) p[k] = k;
#pragma unroll
  for (int j = 0; j < B; ++j) {
    // pivot search and row exchange by predicated selects (constant indices)
    double mx = fabs(M_(j, j));
    int r = j;
#pragma unroll
    for (int k = j + 1; k < B; ++k)
      if (fabs(M_(k, j)) > mx) {
        mx = fabs(M_(k, j));
        r = k;
      }
#pragma unroll
    for (int k = j + 1; k < B; ++k) {
      if (r == k) {
#pragma unroll
        for (int c = 0; c < B; ++c) {
          double t = M_(j, c);
          M_(j, c) = M_(k, c);
          M_(k, c) = t;
        }
        int t = p[j];
        p[j] = p[k];
        p[k] = t;
      }
    }
    if (M_(j, j) == 0.0) return false;
    double hr = 1. / M_(j, j);
    M_(j, j) = hr;
#pragma unroll
    for (int k = 0; k < B; ++k) {
      if (k == j) continue;
#pragma unroll
      for (int a = 0; a < B; ++a) {
        if (a == j) continue;
        M_(a, k) -= M_(a, j) * M_(j, k) * hr;
      }
    }
#pragma unroll
    for (int a = 0; a < B; ++a) {
      M_(a, j) *= hr;
      M_(j, a) *= -hr;
    }
    M_(j, j) = hr;
  }
  // column k of the eliminated matrix belongs to column p[k] of the inverse
  double t[B * B];
#pragma unroll
  for (int a = 0; a < B; ++a)
#pragma unroll
    for (int k = 0; k < B; ++k) t[a * B + p[k]] = M_(a, k);
#pragma unroll
  for (int e = 0; e < B * B; ++e) m[e] = t[e];
#undef M_
  return true;
}

// Inverse diagonal blocks (B x B, row-major, in FP64 or narrowed to DT) of a
// node-block matrix: NodeMatrix::gauss_jordan on each block, the device
// counterpart of invert_device (cuda/cusparsehelper.cu:24-118).  B is a
// template parameter so that every index is a constant and the block lives in
// registers (B*B doubles per lane).
template <int B, typename DT>
__global__ __launch_bounds__(kBlock) void k_block_inverse(int64_t n,
                                                          const int32_t* __restrict__ diagpos,
                                                          const double* __restrict__ val,
                                                          DT* __restrict__ dinv,
                                                          int* __restrict__ singular) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  double m[B * B];
  const double* src = val + (int64_t)diagpos[i] * B * B;
#pragma unroll
  for (int k = 0; k < B * B; ++k) m[k] = src[k];
  if (!gauss_jordan<B>(m)) {
    atomicExch(singular, 1);
    return;
  }
  DT* dst = dinv + i * B * B;
#pragma unroll
  for (int k = 0; k < B * B; ++k) dst[k] = (DT)m[k];
}

// --------------------------------------------------------------- transfers
// uc(i,c) = sum_pos w * uf(col,c)  (restrict_zero); rows in rowmask != 0 are
// written as zero (strong Dirichlet rows folded into the restriction).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_restrict(int64_t nc, int ncomp,
                                                     const int32_t* __restrict__ row,
                                                     const uint32_t* __restrict__ col,
                                                     const T* __restrict__ uf,
                                                     T* __restrict__ uc) {
  const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (t >= nc * ncomp) return;
  const int64_t i = t / ncomp;
  const int c = (int)(t - i * ncomp);
  T acc = (T)0;
  for (int pos = row[i]; pos < row[i + 1]; ++pos) {
    const uint32_t e = col[pos];
    acc = add_rn(acc, mul_rn(code_weight<T>(e >> 29), uf[(int64_t)(e & kIdxMask) * ncomp + c]));
  }
  uc[t] = acc;
}

// uf(i,c) += sum_pos w * uc(col,c)  (prolongate_add, term by term)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_prolongate_add(int64_t nf, int ncomp,
                                                           const int32_t* __restrict__ row,
                                                           const uint32_t* __restrict__ col,
                                                           const T* __restrict__ uc,
                                                           T* __restrict__ uf) {
  const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (t >= nf * ncomp) return;
  const int64_t i = t / ncomp;
  const int c = (int)(t - i * ncomp);
  T acc = uf[t];
  for (int pos = row[i]; pos < row[i + 1]; ++pos) {
    const uint32_t e = col[pos];
    acc = add_rn(acc, mul_rn(code_weight<T>(e >> 29), uc[(int64_t)(e & kIdxMask) * ncomp + c]));
  }
  uf[t] = acc;
}

// mgstep's coarse-grid correction in one pass (stdmultilevelsolver.cc:513-524):
//   v = P u_c (rows of hanging / Dirichlet fine nodes removed = HNZero, SetBoundaryVectorZero)
//   u += mgomega * v
template <typename T, typename UT>
__global__ __launch_bounds__(kBlock) void k_prolongate_correct(int64_t nf, int ncomp,
                                                               const int32_t* __restrict__ row,
                                                               const uint32_t* __restrict__ col,
                                                               const T* __restrict__ uc,
                                                               UT* __restrict__ u, T mgomega) {
  const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (t >= nf * ncomp) return;
  const int64_t i = t / ncomp;
  const int c = (int)(t - i * ncomp);
  const int s = row[i], e = row[i + 1];
  if (s == e) return;
  T v = (T)0;
  for (int pos = s; pos < e; ++pos) {
    const uint32_t q = col[pos];
    v = add_rn(v, mul_rn(code_weight<T>(q >> 29), uc[(int64_t)(q & kIdxMask) * ncomp + c]));
  }
  u[t] = (UT)add_rn((T)u[t], mul_rn(mgomega, v));
}

// -------------------------------------------------------------- hanging nodes
// HNStructure*::Average, parents are regular nodes (checked at creation), so
// every hanging node is independent.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_hn_average(int kind, int64_t ne, const int32_t* __restrict__ edges,
                                                       int64_t nf, const int32_t* __restrict__ faces,
                                                       int ncomp, T* __restrict__ u) {
  const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  const int64_t ntot = (ne + nf) * ncomp;
  if (t >= ntot) return;
  const int64_t k = t / ncomp;
  const int c = (int)(t - k * ncomp);
  const bool q2 = (kind == 22 || kind == 23);
  auto U = [&](int32_t node) { return u[(int64_t)node * ncomp + c]; };
  if (k < ne) {
    const int32_t* e = edges + 4 * k;
    T r;
    if (q2)
      r = add_rn(add_rn(mul_rn((T)0.375, U(e[1])), mul_rn((T)0.75, U(e[2]))),
                 mul_rn((T)-0.125, U(e[3])));
    else
      r = add_rn(mul_rn((T)0.5, U(e[1])), mul_rn((T)0.5, U(e[2])));
    u[(int64_t)e[0] * ncomp + c] = r;
  } else {
    const int32_t* f = faces + 10 * (k - ne);
    const int32_t* g = f + 1;
    T r;
    if (kind == 23) {
      const T w[3] = {(T)0.375, (T)0.75, (T)-0.125};
      T fw[9];
      for (int a = 0; a < 3; ++a)
        for (int b2 = 0; b2 < 3; ++b2) fw[3 * a + b2] = (T)((double)w[a] * (double)w[b2]);
      r = add_rn(add_rn(mul_rn(fw[0], U(g[0])), mul_rn(fw[1], U(g[1]))), mul_rn(fw[2], U(g[2])));
      r = add_rn(r, add_rn(add_rn(mul_rn(fw[3], U(g[3])), mul_rn(fw[4], U(g[4]))),
                           mul_rn(fw[5], U(g[5]))));
      r = add_rn(r, add_rn(add_rn(mul_rn(fw[6], U(g[6])), mul_rn(fw[7], U(g[7]))),
                           mul_rn(fw[8], U(g[8]))));
    } else {
      r = add_rn(add_rn(add_rn(mul_rn((T)0.25, U(g[0])), mul_rn((T)0.25, U(g[1]))),
                        mul_rn((T)0.25, U(g[3]))),
                 mul_rn((T)0.25, U(g[4])));
    }
    u[(int64_t)f[0] * ncomp + c] = r;
  }
}

// HNStructure*::Distribute in gather form: each parent adds its contributions
// in the order the reference's sequential loop would (edges/faces by key).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_hn_gather(int64_t np, int ncomp,
                                                      const int32_t* __restrict__ parent,
                                                      const int32_t* __restrict__ row,
                                                      const int32_t* __restrict__ src,
                                                      const double* __restrict__ w,
                                                      T* __restrict__ u) {
  const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (t >= np * ncomp) return;
  const int64_t k = t / ncomp;
  const int c = (int)(t - k * ncomp);
  const int64_t p = (int64_t)parent[k] * ncomp + c;
  T acc = u[p];
  for (int j = row[k]; j < row[k + 1]; ++j)
    acc = add_rn(acc, mul_rn((T)w[j], u[(int64_t)src[j] * ncomp + c]));
  u[p] = acc;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_zero_nodes(int64_t m, int ncomp,
                                                       const int32_t* __restrict__ nodes,
                                                       T* __restrict__ u) {
  const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (t >= m * ncomp) return;
  const int64_t k = t / ncomp;
  u[(int64_t)nodes[k] * ncomp + (t - k * ncomp)] = (T)0;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_zero_flat(int64_t m, const int32_t* __restrict__ idx,
                                                      T* __restrict__ u) {
  const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (t >= m) return;
  u[idx[t]] = (T)0;
}

// ------------------------------------------------------- dense coarse solve
// x = Ainv * b with the FP64 inverse of a direct level; one wavefront per row.
template <typename BT, typename XT>
__global__ __launch_bounds__(kBlock) void k_dense_solve(int64_t m, const double* __restrict__ Ainv,
                                                        const BT* __restrict__ b,
                                                        XT* __restrict__ x) {
  const int64_t r = (int64_t)blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (r >= m) return;
  double acc = 0.0;
  for (int64_t c = lane; c < m; c += 64) acc += Ainv[r * m + c] * (double)b[c];
  acc = wave_sum(acc);
  if (lane == 0) x[r] = (XT)acc;
}

}  // namespace gh
