// Vanka patch smoother (src/LinAlg/vankasmoother.cc): the patch matrices of a
// node-block matrix, LU with partial pivoting on the device, and the solve
//   x <- (1/count_i) sum_p R_p^T A_p^-1 R_p x.
// Factoring: one workgroup of 1024 threads per patch, the m x m patch matrix
// (m <= 128) in the threads' registers, the LU's operations uncontracted so
// that the factors equal the oracle's; the recorded row exchanges are composed on the host into
// one gather per factor row.  Solving: one wavefront per patch, the local vector
// gathered in pivoted order into the lanes' registers (rows lane and
// lane + 64), column-oriented triangular solves (each pass reads only its
// triangle) with the solved values broadcast by readlane.  Averaging: one
// thread per unknown, the patches containing its node added in ascending order
// (the reference adds with OpenMP atomics, in no fixed order).
#include <algorithm>
#include <cmath>
#include <vector>

#include "gh_kernels.h"

using namespace gh;

namespace {

constexpr int kVankaMaxM = 128;       // local unknowns per patch
constexpr int kVankaMaxNodes = 128;   // nodes per patch

// Factoring one patch: 1024 threads, lane tr = t % 64 of wavefront tc = t / 64
// keeps the entries (tr + 64 i, tc + 16 j), i < 2, j < 8, of the m x m patch
// matrix (m <= 128) in registers, so the 64 owners of a column form one
// wavefront and its pivot search is a wavefront reduction.  Row exchanges are
// not carried out: every thread keeps the position of each of its rows in the
// exchanged order, and the factors are written in that order at the end.
// LDS holds only the patch's CSR positions and, per column, the multipliers
// and the pivot (double-buffered), a few KB, so two patches share a CU.
constexpr int kFactorThreads = 1024;
constexpr int kFactorLanes = 64;                          // rows tr + 64 i
constexpr int kFactorWaves = kFactorThreads / kFactorLanes;  // columns tc + 16 j
constexpr int kRowsPer = kVankaMaxM / kFactorLanes;       // 2
constexpr int kColsPer = kVankaMaxM / kFactorWaves;       // 8

// DPP move of a double (both halves); lanes the masks leave out read 0.
template <int CTRL, int ROWMASK, int BANKMASK>
__device__ __forceinline__ double dpp_move(double x) {
  const int64_t b = __builtin_bit_cast(int64_t, x);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xFFFFFFFF), CTRL, ROWMASK, BANKMASK, true);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, ROWMASK, BANKMASK, true);
  return __builtin_bit_cast(double, ((int64_t)hi << 32) | (uint32_t)lo);
}

__device__ __forceinline__ double lane_read(double x, int lane) {
  const int64_t b = __builtin_bit_cast(int64_t, x);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xFFFFFFFF), lane);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
  return __builtin_bit_cast(double, ((int64_t)hi << 32) | (uint32_t)lo);
}

// Maximum over the 64 lanes of a wavefront of values >= 0 (or -1 for none):
// row shifts inside each row of 16 lanes, then the row broadcasts; lane 63
// ends with the maximum.
__device__ __forceinline__ double wave_max_nonneg(double x) {
  x = fmax(x, dpp_move<0x111, 0xf, 0xf>(x));  // row_shr:1
  x = fmax(x, dpp_move<0x112, 0xf, 0xf>(x));  // row_shr:2
  x = fmax(x, dpp_move<0x114, 0xf, 0xe>(x));  // row_shr:4
  x = fmax(x, dpp_move<0x118, 0xf, 0xc>(x));  // row_shr:8
  x = fmax(x, dpp_move<0x142, 0xa, 0xf>(x));  // row_bcast:15
  x = fmax(x, dpp_move<0x143, 0xc, 0xf>(x));  // row_bcast:31
  return lane_read(x, 63);
}

// Minimum over the 64 lanes of a wavefront, the same pattern; lanes the masks
// leave out keep INT_MAX.
template <int CTRL, int ROWMASK, int BANKMASK>
__device__ __forceinline__ int dpp_min(int x) {
  return min(x, __builtin_amdgcn_update_dpp(INT_MAX, x, CTRL, ROWMASK, BANKMASK, false));
}

__device__ __forceinline__ int wave_min(int x) {
  x = dpp_min<0x111, 0xf, 0xf>(x);
  x = dpp_min<0x112, 0xf, 0xf>(x);
  x = dpp_min<0x114, 0xf, 0xe>(x);
  x = dpp_min<0x118, 0xf, 0xc>(x);
  x = dpp_min<0x142, 0xa, 0xf>(x);
  x = dpp_min<0x143, 0xc, 0xf>(x);
  return __builtin_amdgcn_readlane(x, 63);
}

struct FactorShared {
  double l[2][kVankaMaxM];  // the multipliers of the column, by row
  int32_t pr[2], rp[2];     // the pivot's position before the exchange, and its row
  double d[2];              // the pivot
};

// Columns 16 KB .. 16 KB + 15 (KB a constant, so that the column's registers
// are known): per column k, one barrier.
//   A: the wavefront owning column k (already updated by itself) finds the
//      pivot by a maximum of |a_ik| and a minimum of position * 256 + row,
//      forms the multipliers, puts them and the pivot into LDS, records
//      piv[k];
//   B: every thread exchanges the positions k and piv[k] among its rows and
//      updates its entries below and right of (k, k): the pivot row's entries
//      in a wavefront's columns are held by one of its own lanes (rp % 64),
//      which broadcasts them.
// The buffers alternate between columns, so the owner of column k + 1 may
// start before the others have finished column k.
template <int KB>
__device__ __forceinline__ void factor_columns(int m, int tr, int tc,
                                               double (&a)[kRowsPer][kColsPer],
                                               int (&pos)[kRowsPer], FactorShared& S,
                                               int32_t* __restrict__ pv) {
  for (int kk = 0; kk < kFactorWaves; ++kk) {
    const int k = KB * kFactorWaves + kk;
    if (k >= m) return;
    const int buf = k & 1;
    if (tc == kk) {
      double best = -1.0;  // fmax passes NaN over
#pragma unroll
      for (int i = 0; i < kRowsPer; ++i)
        if (pos[i] >= k) best = fmax(best, fabs(a[i][KB]));
      best = wave_max_nonneg(best);
      // the smallest position holding the maximum (a NaN a_kk keeps row k, as
      // the oracle's strict comparisons do); a key always exists, since a
      // non-NaN a_kk makes best >= |a_kk| attained
      int key = INT_MAX;
#pragma unroll
      for (int i = 0; i < kRowsPer; ++i) {
        const double v = a[i][KB];
        const bool cand = pos[i] > k ? fabs(v) == best
                                     : pos[i] == k && (fabs(v) == best || v != v);
        if (cand) key = min(key, pos[i] * 256 + tr + kFactorLanes * i);
      }
      key = wave_min(key);
      const int pr = key >> 8, rp = key & 255;
      const double d = lane_read(rp < kFactorLanes ? a[0][KB] : a[1][KB], rp % kFactorLanes);
      if (d != 0.0) {
#pragma unroll
        for (int i = 0; i < kRowsPer; ++i) {
          const int r = tr + kFactorLanes * i;
          if (pos[i] >= k && r != rp) {
            const double l = a[i][KB] / d;
            a[i][KB] = l;
            S.l[buf][r] = l;
          }
        }
      }
      if (tr == 0) {
        S.pr[buf] = pr;
        S.rp[buf] = rp;
        S.d[buf] = d;
        pv[k] = pr;
      }
    }
    __syncthreads();
    const int pr = __builtin_amdgcn_readfirstlane(S.pr[buf]);
    const int rp = __builtin_amdgcn_readfirstlane(S.rp[buf]);
    const double d = S.d[buf];
#pragma unroll
    for (int i = 0; i < kRowsPer; ++i) {
      const int r = tr + kFactorLanes * i;
      pos[i] = (r == rp) ? k : (pos[i] == k ? pr : pos[i]);
    }
    if (d != 0.0) {
      double l[kRowsPer];
#pragma unroll
      for (int i = 0; i < kRowsPer; ++i) l[i] = pos[i] > k ? S.l[buf][tr + kFactorLanes * i] : 0.0;
#pragma unroll
      for (int j = KB; j < kColsPer; ++j) {
        // column tc + 16 j > k: every j > KB, and j = KB right of the pivot
        if (!((j > KB || tc > kk) && tc + kFactorWaves * j < m)) continue;
        // the pivot row's entry in this column, from the lane holding it
        const double u = lane_read(rp < kFactorLanes ? a[0][j] : a[1][j], rp % kFactorLanes);
#pragma unroll
        for (int i = 0; i < kRowsPer; ++i)
          if (pos[i] > k) a[i][j] = add_rn(a[i][j], -mul_rn(l[i], u));
      }
    }
  }
}

// LU with partial pivoting of every patch matrix, the oracle's operations in
// the oracle's order per entry: pivot = the first row (in the exchanged order)
// of largest |a_ik|, i >= k; no elimination when the pivot is 0;
// l_i = a_ik / d and a_ij -= l_i a_kj, product and difference rounded
// separately.  (8 waves per SIMD: two patches per CU, at most 64 VGPRs.)
__global__ __launch_bounds__(kFactorThreads, 8) void k_vanka_factor(
    int64_t npatches, int psize, int B, const int32_t* __restrict__ nodes,
    const int32_t* __restrict__ rowptr, const int32_t* __restrict__ col,
    const double* __restrict__ val, double* __restrict__ lu64, float* __restrict__ lu32,
    int32_t* __restrict__ piv) {
  extern __shared__ int32_t tab[];  // psize * psize CSR positions of the node blocks, -1: none
  __shared__ int32_t pn[kVankaMaxNodes];
  __shared__ FactorShared S;
  constexpr int NT = kFactorThreads;
  const int64_t p = blockIdx.x;
  const int m = psize * B;
  const int t = threadIdx.x, tr = t % kFactorLanes, tc = t / kFactorLanes;
  for (int e = t; e < psize * psize; e += NT) tab[e] = -1;
  for (int r = t; r < psize; r += NT) pn[r] = nodes[p * psize + r];
  __syncthreads();
  // copy_entries_sparseblockmatrix: the entries whose column is a patch node.
  // NT / psize threads (at least 8) walk each row's entries.
  {
    const int per = NT / psize;
    const int r = t / per, u0 = t - r * per;
    if (r < psize) {
      const int32_t row = pn[r];
      for (int32_t q = rowptr[row] + u0; q < rowptr[row + 1]; q += per) {
        const int32_t c0 = col[q];
        for (int u = 0; u < psize; ++u)
          if (pn[u] == c0) tab[r * psize + u] = q;
      }
    }
  }
  __syncthreads();
  double a[kRowsPer][kColsPer];
  int pos[kRowsPer];  // position of row tr + 64 i in the exchanged order; -1: no row
#pragma unroll
  for (int i = 0; i < kRowsPer; ++i) {
    const int r = tr + kFactorLanes * i;
    pos[i] = r < m ? r : -1;
#pragma unroll
    for (int j = 0; j < kColsPer; ++j) {
      const int c = tc + kFactorWaves * j;
      double x = 0.0;
      if (r < m && c < m) {
        const int32_t q = tab[(r / B) * psize + c / B];
        if (q >= 0) x = val[((int64_t)q * B + r % B) * B + c % B];
      }
      a[i][j] = x;
    }
  }
  int32_t* pv = piv + p * m;
  factor_columns<0>(m, tr, tc, a, pos, S, pv);
  factor_columns<1>(m, tr, tc, a, pos, S, pv);
  factor_columns<2>(m, tr, tc, a, pos, S, pv);
  factor_columns<3>(m, tr, tc, a, pos, S, pv);
  factor_columns<4>(m, tr, tc, a, pos, S, pv);
  factor_columns<5>(m, tr, tc, a, pos, S, pv);
  factor_columns<6>(m, tr, tc, a, pos, S, pv);
  factor_columns<7>(m, tr, tc, a, pos, S, pv);
  // column-major factors (coalesced column reads in the solve), rows in the
  // exchanged order
  const int64_t base = p * (int64_t)m * m;
#pragma unroll
  for (int i = 0; i < kRowsPer; ++i) {
    if (pos[i] < 0) continue;
#pragma unroll
    for (int j = 0; j < kColsPer; ++j) {
      const int c = tc + kFactorWaves * j;
      if (c >= m) continue;
      const int64_t e = base + (int64_t)c * m + pos[i];
      if (lu64) lu64[e] = a[i][j];
      else lu32[e] = (float)a[i][j];
    }
  }
}

template <typename T>
__device__ __forceinline__ T bcast(T v, int lane) {
  if constexpr (sizeof(T) == 8) {
    const int64_t b = __builtin_bit_cast(int64_t, v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xFFFFFFFF), lane);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
    return __builtin_bit_cast(T, ((int64_t)hi << 32) | (uint32_t)lo);
  } else {
    return __builtin_bit_cast(T, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), lane));
  }
}

// one wavefront per patch; rows lane (slot 0) and lane + 64 (slot 1)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_vanka_solve(int64_t npatches, int psize, int B,
                                                        const T* __restrict__ lu,
                                                        const int32_t* __restrict__ gidx,
                                                        const T* __restrict__ h,
                                                        T* __restrict__ xp) {
  const int lane = threadIdx.x & 63;
  const int64_t p = (int64_t)blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
  if (p >= npatches) return;  // the whole wavefront
  const int m = psize * B;
  const T* L = lu + p * (int64_t)m * m;
  // H = P H: gidx holds, per row of the factors, the global entry that the
  // recorded row exchanges bring there
  const int32_t* gi = gidx + p * m;
  T v0 = (T)0, v1 = (T)0;
  if (lane < m) v0 = h[gi[lane]];
  if (lane + 64 < m) v1 = h[gi[lane + 64]];
  // The column loads of a batch of kBatch steps are issued together, before
  // the batch's (sequential) updates; every element still sees the same
  // operations in the same order.  Each pass reads only its triangle: the
  // forward pass the rows below the diagonal, the backward pass the rows
  // above it and the diagonal.  (Measured on 32768 patches of 108 unknowns:
  // batches of 16 beat 8 and 32, and beat double-buffered batches, whose
  // register cost halves the wavefronts in flight.)
  constexpr int kBatch = 16;
  const bool in0 = lane < m, in1 = lane + 64 < m;
  // unit lower: column k updates rows i > k
  for (int k0 = 0; k0 < m - 1; k0 += kBatch) {
    T a0[kBatch], a1[kBatch];
#pragma unroll
    for (int u = 0; u < kBatch; ++u) {
      const int k = min(k0 + u, m - 1);
      const T* Lk = L + (int64_t)k * m;
      a0[u] = (in0 && lane > k) ? Lk[lane] : (T)0;
      a1[u] = (in1 && lane + 64 > k) ? Lk[lane + 64] : (T)0;
    }
#pragma unroll
    for (int u = 0; u < kBatch; ++u) {
      const int k = k0 + u;
      if (k < m - 1) {
        const T yk = bcast(k < 64 ? v0 : v1, k & 63);
        if (lane > k && in0) v0 = add_rn(v0, -mul_rn(a0[u], yk));
        if (lane + 64 > k && in1) v1 = add_rn(v1, -mul_rn(a1[u], yk));
      }
    }
  }
  // upper: h_k /= u_kk, then column k updates rows i < k
  for (int k0 = m - 1; k0 >= 0; k0 -= kBatch) {
    T a0[kBatch], a1[kBatch], dg[kBatch];
#pragma unroll
    for (int u = 0; u < kBatch; ++u) {
      const int kk = max(k0 - u, 0);
      const T* Uk = L + (int64_t)kk * m;
      a0[u] = lane < kk ? Uk[lane] : (T)0;
      a1[u] = lane + 64 < kk ? Uk[lane + 64] : (T)0;
      dg[u] = Uk[kk];
    }
#pragma unroll
    for (int u = 0; u < kBatch; ++u) {
      const int k = k0 - u;
      if (k >= 0) {
        if (lane == (k & 63)) {
          if (k < 64) v0 = v0 / dg[u];
          else v1 = v1 / dg[u];
        }
        const T yk = bcast(k < 64 ? v0 : v1, k & 63);
        if (lane < k) v0 = add_rn(v0, -mul_rn(a0[u], yk));
        if (lane + 64 < k) v1 = add_rn(v1, -mul_rn(a1[u], yk));
      }
    }
  }
  T* out = xp + p * (int64_t)m;
  if (lane < m) out[lane] = v0;
  if (lane + 64 < m) out[lane + 64] = v1;
}

// x(i, c) = (sum over the patches of node i, ascending) * (1 / count_i)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_vanka_average(int64_t n, int B, int psize, int m,
                                                          const int32_t* __restrict__ nptr,
                                                          const int32_t* __restrict__ nent,
                                                          const T* __restrict__ xp,
                                                          T* __restrict__ x) {
  const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (t >= n * B) return;
  const int64_t i = t / B;
  const int c = (int)(t - i * B);
  const int32_t e0 = nptr[i], e1 = nptr[i + 1];
  T acc = (T)0;
  for (int32_t e = e0; e < e1; ++e) {
    const int32_t pe = nent[e];
    const int64_t pp = pe / psize;
    const int r = pe - (int)(pp * psize);
    acc = add_rn(acc, xp[pp * m + B * r + c]);
  }
  const double s = 1.0 / (double)(e1 - e0);
  x[t] = (T)mul_rn((double)acc, s);
}

}  // namespace

namespace gh {

int vanka_solve(gh_ctx* ctx, const gh_vanka* V, int prec, void* x) {
  if (prec != V->prec) return fail(GH_EINVAL, "vanka: vector precision differs from the factors'");
  const unsigned gs = (unsigned)((V->npatches + kBlock / 64 - 1) / (kBlock / 64));
  const unsigned ga = (unsigned)((V->n * V->B + kBlock - 1) / kBlock);
  if (V->npatches == 0) return GH_OK;
  if (prec == GH_F64) {
    hipLaunchKernelGGL(k_vanka_solve<double>, dim3(gs), dim3(kBlock), 0, ctx->stream, V->npatches,
                       V->psize, V->B, (const double*)V->lu, V->piv, (const double*)x,
                       (double*)V->xp);
    hipLaunchKernelGGL(k_vanka_average<double>, dim3(ga), dim3(kBlock), 0, ctx->stream, V->n, V->B,
                       V->psize, V->m, V->nptr, V->nent, (const double*)V->xp, (double*)x);
  } else {
    hipLaunchKernelGGL(k_vanka_solve<float>, dim3(gs), dim3(kBlock), 0, ctx->stream, V->npatches,
                       V->psize, V->B, (const float*)V->lu, V->piv, (const float*)x,
                       (float*)V->xp);
    hipLaunchKernelGGL(k_vanka_average<float>, dim3(ga), dim3(kBlock), 0, ctx->stream, V->n, V->B,
                       V->psize, V->m, V->nptr, V->nent, (const float*)V->xp, (float*)x);
  }
  GH_HIP(hipGetLastError());
  return GH_OK;
}

int vanka_create(gh_ctx* ctx, const gh_csr* A, int64_t npatches, int psize, const int32_t* nodes,
                 int prec, gh_vanka** out) {
  GH_ARG(ctx && A && out && (nodes || npatches == 0) && npatches >= 0 && psize >= 1,
         "gh_vanka_create: bad arguments");
  GH_ARG(A->prec == GH_F64, "gh_vanka_create: the matrix must be FP64");
  GH_ARG(prec == GH_F64 || prec == GH_F32, "gh_vanka_create: factors are FP64 or FP32");
  const int B = A->ncomp, m = psize * B;
  GH_ARG(psize <= kVankaMaxNodes && m <= kVankaMaxM,
         "gh_vanka_create: at most 128 unknowns per patch");
  GH_ARG(A->nnz > 0 || npatches == 0, "gh_vanka_create: patches of a matrix without entries");
  const int64_t n = A->n;
  // node -> (patch, local) lists, ascending patch; every node must be covered
  std::vector<int32_t> nptr(n + 1, 0), nent((size_t)npatches * psize);
  for (int64_t k = 0; k < npatches * psize; ++k) {
    if (nodes[k] < 0 || nodes[k] >= n) return fail(GH_EINVAL, "gh_vanka_create: node out of range");
    nptr[nodes[k] + 1]++;
  }
  for (int64_t p = 0; p < npatches; ++p)
    for (int r = 0; r < psize; ++r)
      for (int t = 0; t < r; ++t)
        if (nodes[p * psize + t] == nodes[p * psize + r])
          return fail(GH_EINVAL, "gh_vanka_create: patch %lld lists a node twice", (long long)p);
  for (int64_t i = 0; i < n; ++i) {
    if (nptr[i + 1] == 0) return fail(GH_EINVAL, "gh_vanka_create: node %lld is in no patch", (long long)i);
    nptr[i + 1] += nptr[i];
  }
  {
    std::vector<int32_t> fill(nptr.begin(), nptr.end() - 1);
    for (int64_t p = 0; p < npatches; ++p)
      for (int r = 0; r < psize; ++r) nent[fill[nodes[p * psize + r]]++] = (int32_t)(p * psize + r);
  }
  gh_vanka* V = new gh_vanka();
  V->ctx = ctx;
  ctx_retain(ctx);
  V->B = B;
  V->psize = psize;
  V->m = m;
  V->prec = prec;
  V->n = n;
  V->npatches = npatches;
  auto bail = [&](int rc) {
    gh_vanka_destroy(V);
    return rc;
  };
  const size_t ts = prec == GH_F64 ? 8 : 4;
  auto alloc = [&](void** p, size_t bytes) {
    return hipMalloc(p, std::max<size_t>(bytes, 8)) == hipSuccess;
  };
  if (!alloc((void**)&V->nodes, sizeof(int32_t) * npatches * psize) ||
      !alloc(&V->lu, ts * npatches * m * m) || !alloc((void**)&V->piv, sizeof(int32_t) * npatches * m) ||
      !alloc((void**)&V->nptr, sizeof(int32_t) * (n + 1)) ||
      !alloc((void**)&V->nent, sizeof(int32_t) * npatches * psize) ||
      !alloc(&V->xp, ts * npatches * m))
    return bail(fail(GH_ENOMEM, "gh_vanka_create: allocation failed"));
  if (hipMemcpy(V->nodes, nodes, sizeof(int32_t) * npatches * psize, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(V->nptr, nptr.data(), sizeof(int32_t) * (n + 1), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(V->nent, nent.data(), sizeof(int32_t) * npatches * psize, hipMemcpyHostToDevice) != hipSuccess)
    return bail(fail(GH_EHIP, "gh_vanka_create: upload failed"));
  if (npatches > 0) {
    const size_t lds = sizeof(int32_t) * (size_t)psize * psize;
    if (hipFuncSetAttribute((const void*)k_vanka_factor, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)lds) != hipSuccess)
      return bail(fail(GH_EHIP, "gh_vanka_create: %zu bytes of LDS per patch not available", lds));
    hipLaunchKernelGGL(k_vanka_factor, dim3((unsigned)npatches), dim3(kFactorThreads), lds,
                       ctx->stream,
                       npatches, psize, B, V->nodes, A->rowptr, A->col, (const double*)A->val,
                       prec == GH_F64 ? (double*)V->lu : nullptr,
                       prec == GH_F64 ? nullptr : (float*)V->lu, V->piv);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(ctx->stream) != hipSuccess)
      return bail(fail(GH_EHIP, "gh_vanka_create: factorisation failed"));
    // Compose the recorded row exchanges (applied in order k = 0..m-1) into
    // one gather: row i of P h is local entry src[i], i.e. the global entry
    // node(src[i] / B) * B + src[i] % B.  The solve then needs no exchanges.
    std::vector<int32_t> piv((size_t)npatches * m);
    if (hipMemcpy(piv.data(), V->piv, sizeof(int32_t) * piv.size(), hipMemcpyDeviceToHost) != hipSuccess)
      return bail(fail(GH_EHIP, "gh_vanka_create: pivot download failed"));
    std::vector<int32_t> src(m);
    for (int64_t p = 0; p < npatches; ++p) {
      int32_t* pv = piv.data() + p * m;
      for (int i = 0; i < m; ++i) src[i] = i;
      for (int k = 0; k < m; ++k) {
        if (pv[k] < k || pv[k] >= m) return bail(fail(GH_EINVAL, "gh_vanka_create: bad pivot"));
        std::swap(src[k], src[pv[k]]);
      }
      for (int i = 0; i < m; ++i)
        pv[i] = (int32_t)((int64_t)nodes[p * psize + src[i] / B] * B + src[i] % B);
    }
    if (hipMemcpy(V->piv, piv.data(), sizeof(int32_t) * piv.size(), hipMemcpyHostToDevice) != hipSuccess)
      return bail(fail(GH_EHIP, "gh_vanka_create: gather upload failed"));
  }
  *out = V;
  return GH_OK;
}

}  // namespace gh

extern "C" {

int gh_vanka_create(gh_ctx* ctx, const gh_csr* A, int64_t npatches, int psize,
                    const int32_t* nodes, int prec, gh_vanka** out) {
  GH_ARG(A, "gh_vanka_create: NULL matrix");
  GH_ARG(!A->host_only, "gh_vanka_create: the matrix has no device layout (created by gh_csr_create_host)");
  return vanka_create(ctx, A, npatches, psize, nodes, prec, out);
}

int gh_vanka_destroy(gh_vanka* V) {
  if (!V) return GH_OK;
  if (V->ctx && V->ctx->stream) (void)hipStreamSynchronize(V->ctx->stream);
  for (void* p : {(void*)V->nodes, V->lu, (void*)V->piv, (void*)V->nptr, (void*)V->nent, V->xp})
    if (p) (void)hipFree(p);
  gh_ctx* ctx = V->ctx;
  delete V;
  if (ctx) ctx_release(ctx);
  return GH_OK;
}

int gh_vanka_solve(gh_ctx* ctx, const gh_vanka* V, gh_vec* x) {
  GH_ARG(ctx && V && x, "gh_vanka_solve: NULL argument");
  GH_ARG(x->n == V->n && x->ncomp == V->B, "gh_vanka_solve: vector shape does not match");
  return vanka_solve(ctx, V, x->prec, x->d);
}

}  // extern "C"
