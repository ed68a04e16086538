// Vanka patch smoother (src/LinAlg/vankasmoother.cc): the patch matrices of a
// node-block matrix, LU with partial pivoting on the device, and the solve
//   x <- (1/count_i) sum_p R_p^T A_p^-1 R_p x.
// Factoring: one workgroup of 1024 threads per patch, the m x m patch matrix
// (m <= 128) in LDS, the LU's operations uncontracted so that the factors
// equal the oracle's; the recorded row exchanges are composed on the host into
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

// Factoring one patch: the m x m matrix in LDS (93 KB at m = 108, so one
// patch per CU), 1024 threads, two barriers per column (below).
constexpr int kFactorThreads = 1024;

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

__global__ __launch_bounds__(kFactorThreads) void k_vanka_factor(
    int64_t npatches, int psize, int B, const int32_t* __restrict__ nodes,
    const int32_t* __restrict__ rowptr, const int32_t* __restrict__ col,
    const double* __restrict__ val, double* __restrict__ lu64, float* __restrict__ lu32,
    int32_t* __restrict__ piv) {
  extern __shared__ double a[];  // m * m, row-major
  __shared__ int32_t pn[kVankaMaxNodes];
  __shared__ int s_piv;
  __shared__ double s_pv, s_akk;
  constexpr int NT = kFactorThreads;
  const int64_t p = blockIdx.x;
  const int m = psize * B;
  const int t = threadIdx.x;
  for (int e = t; e < m * m; e += NT) a[e] = 0.0;
  for (int r = t; r < psize; r += NT) pn[r] = nodes[p * psize + r];
  __syncthreads();
  // copy_entries_sparseblockmatrix: entries whose column is a patch node.
  // The patch's rows share the threads: NT / psize threads walk each row's
  // entries (every entry lands in its own place, so the order is free).
  {
    const int per = NT / psize;  // psize <= 128, so at least 8 threads per row
    const int r = t / per, u0 = t - r * per;
    if (r < psize) {
      const int32_t row = pn[r];
      for (int32_t q = rowptr[row] + u0; q < rowptr[row + 1]; q += per) {
        const int32_t c0 = col[q];
        int c = -1;
        for (int u = 0; u < psize; ++u)
          if (pn[u] == c0) c = u;
        if (c < 0) continue;
        for (int cr = 0; cr < B; ++cr)
          for (int cc = 0; cc < B; ++cc)
            a[(B * r + cr) * m + B * c + cc] = val[((int64_t)q * B + cr) * B + cc];
      }
    }
  }
  __syncthreads();
  // LU with partial pivoting (first largest |a_ik|), two barriers per column:
  //   phase 1: swap rows k and pr (every column but k), divide column k by the
  //            pivot d (the division of row pr takes the old a_kk, saved by
  //            the search, since the swap moves that row's column-k entry);
  //   phase 2: a_ij -= l_i a_kj; wavefront 0 updates column k + 1 itself and
  //            searches it for the next pivot in the same phase, the others
  //            update columns k + 2, ... (every product and difference
  //            rounded separately, as the oracle computes them).
  // The search (wavefront 0; lanes hold rows k + lane and k + lane + 64): the
  // largest |a_ik| by a DPP maximum, then the first row holding it (rows
  // k..k+63 before k+64..) by ballots.
  auto search = [&](int k, double v0, double v1) {
    const bool in0 = k + t < m, in1 = k + t + 64 < m;
    const double best = wave_max_nonneg(fmax(in0 ? fabs(v0) : -1.0, in1 ? fabs(v1) : -1.0));
    const uint64_t b0 = __ballot(in0 && fabs(v0) == best);
    int bi;
    double bv;
    if (b0) {
      const int l = __builtin_ctzll(b0);
      bi = k + l;
      bv = lane_read(v0, l);
    } else {
      const uint64_t b1 = __ballot(in1 && fabs(v1) == best);
      const int l = __builtin_ctzll(b1);
      bi = k + 64 + l;
      bv = lane_read(v1, l);
    }
    const double akk = lane_read(v0, 0);  // row k itself
    if (t == 0) {
      s_piv = bi;
      s_pv = bv;
      s_akk = akk;
    }
  };
  if (t < 64) search(0, t < m ? a[t * m] : 0.0, t + 64 < m ? a[(t + 64) * m] : 0.0);
  __syncthreads();
  constexpr int kUpd = NT - 64;            // threads updating columns k + 2 ...
  constexpr int kGroups = kUpd / 128;      // 7 row groups of 128 columns
  const int tu = t - 64, tj = tu & 127, tg = tu >> 7;
  for (int k = 0; k < m; ++k) {
    const int pr = s_piv;
    const double d = s_pv, akk = s_akk;
    if (pr != k)
      for (int j = t; j < m; j += NT) {
        if (j == k) continue;
        const double x = a[k * m + j];
        a[k * m + j] = a[pr * m + j];
        a[pr * m + j] = x;
      }
    if (t == 0) {
      a[k * m + k] = (pr != k) ? d : akk;
      piv[p * m + k] = pr;
    }
    if (d != 0.0) {
      for (int i = k + 1 + t; i < m; i += NT) {
        const double x = (i == pr) ? akk : a[i * m + k];
        a[i * m + k] = x / d;
      }
    } else if (pr != k && t == 0) {
      a[pr * m + k] = akk;
    }
    __syncthreads();
    if (t < 64) {
      // column k + 1: update (when d != 0), then search it
      const int j = k + 1;
      if (j < m) {
        double v0 = 0.0, v1 = 0.0;
        const double akj = a[k * m + j];
        const int i0 = j + t, i1 = j + t + 64;
        if (i0 < m) {
          v0 = a[i0 * m + j];
          if (d != 0.0) {
            v0 = add_rn(v0, -mul_rn(a[i0 * m + k], akj));
            a[i0 * m + j] = v0;
          }
        }
        if (i1 < m) {
          v1 = a[i1 * m + j];
          if (d != 0.0) {
            v1 = add_rn(v1, -mul_rn(a[i1 * m + k], akj));
            a[i1 * m + j] = v1;
          }
        }
        search(j, v0, v1);
      }
    } else if (d != 0.0 && tg < kGroups) {
      const int j = k + 2 + tj;
      if (j < m) {
        const double akj = a[k * m + j];
        int i = k + 1 + tg;
        for (; i + kGroups < m; i += 2 * kGroups) {
          const double l0 = a[i * m + k], l1 = a[(i + kGroups) * m + k];
          const double x0 = a[i * m + j], x1 = a[(i + kGroups) * m + j];
          a[i * m + j] = add_rn(x0, -mul_rn(l0, akj));
          a[(i + kGroups) * m + j] = add_rn(x1, -mul_rn(l1, akj));
        }
        if (i < m) a[i * m + j] = add_rn(a[i * m + j], -mul_rn(a[i * m + k], akj));
      }
    }
    __syncthreads();
  }
  // column-major factors for coalesced column reads in the solve
  const int64_t base = p * (int64_t)m * m;
  for (int e = t; e < m * m; e += NT) {
    const int k = e / m, i = e - k * m;
    if (lu64) lu64[base + e] = a[i * m + k];
    else lu32[base + e] = (float)a[i * m + k];
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
    const size_t lds = sizeof(double) * (size_t)m * m;
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
