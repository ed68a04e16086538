// Runtime of the gascoignelib_amd C ABI: context, vectors, matrices, grid
// transfer, hanging nodes, Dirichlet masks and the launch wrappers around the
// kernels of gh_kernels.h.
#include <algorithm>
#include <type_traits>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <unordered_map>

#include "gh_sell.h"

// Scratch of gh_jacobi_noalloc: 8 MiB per GPU, held until the process ends.
// Host code never passes `scratch` itself to a kernel (in host code the name
// stands for a host-side shadow, not a device address): gh_ctx_create looks up
// each device's address with hipGetSymbolAddress, and only that address is used.
static constexpr size_t kNoAllocScratchBytes = size_t(8) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];
static constexpr int kMaxDevices = 64;
static std::mutex g_noalloc_mu;
static void* g_noalloc_ptr[kMaxDevices] = {};

namespace gh {

static thread_local std::string g_err;

void set_error(const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
}

int fail(int code, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

static size_t prec_bytes(int prec) { return prec == GH_F64 ? 8 : prec == GH_F32 ? 4 : 2; }

static unsigned grid_stride(int64_t n) {
  int64_t g = (n + kBlock - 1) / kBlock;
  if (g > 8192) g = 8192;
  if (g < 1) g = 1;
  return (unsigned)g;
}

int vec_alloc(gh_ctx* ctx, int64_t n, int ncomp, int prec, gh_vec* v) {
  if (n < 0 || ncomp < 1 || (prec != GH_F64 && prec != GH_F32))
    return fail(GH_EINVAL, "gh_vec_create: bad shape n=%lld ncomp=%d prec=%d", (long long)n,
                ncomp, prec);
  v->ctx = ctx;
  v->n = n;
  v->ncomp = ncomp;
  v->prec = prec;
  v->size = n * ncomp;
  v->d = nullptr;
  size_t bytes = (size_t)std::max<int64_t>(v->size, 1) * prec_bytes(prec);
  if (hipMalloc(&v->d, bytes) != hipSuccess)
    return fail(GH_ENOMEM, "gh_vec_create: hipMalloc(%zu) failed", bytes);
  GH_HIP(hipMemsetAsync(v->d, 0, bytes, ctx->stream));
  return GH_OK;
}

void vec_free(gh_vec* v) {
  if (v->d && v->owned) (void)hipFree(v->d);
  v->d = nullptr;
}

int fetch_scalars(gh_ctx* ctx, int first, int count, double* out) {
  GH_HIP(hipMemcpyAsync(ctx->host_scal + first, ctx->scal + first, sizeof(double) * count,
                        hipMemcpyDeviceToHost, ctx->stream));
  GH_HIP(hipStreamSynchronize(ctx->stream));
  for (int i = 0; i < count; ++i) out[i] = ctx->host_scal[first + i];
  return GH_OK;
}

int finalize_slots(gh_ctx* ctx, int first, int count, bool max) {
  if (max)
    hipLaunchKernelGGL(k_finalize<true>, dim3(count), dim3(kBlock), 0, ctx->stream,
                       ctx->partials + (int64_t)first * kRedBlocks, ctx->scal + first);
  else
    hipLaunchKernelGGL(k_finalize<false>, dim3(count), dim3(kBlock), 0, ctx->stream,
                       ctx->partials + (int64_t)first * kRedBlocks, ctx->scal + first);
  GH_HIP(hipGetLastError());
  return GH_OK;
}

int reduce_dot(gh_ctx* ctx, int prec, int64_t size, const void* x, const void* y, int slot) {
  double* part = ctx->partials + (int64_t)slot * kRedBlocks;
  if (prec == GH_F64)
    hipLaunchKernelGGL(k_dot_part<double>, dim3(kRedBlocks), dim3(kBlock), 0, ctx->stream, size,
                       (const double*)x, (const double*)y, part);
  else
    hipLaunchKernelGGL(k_dot_part<float>, dim3(kRedBlocks), dim3(kBlock), 0, ctx->stream, size,
                       (const float*)x, (const float*)y, part);
  GH_HIP(hipGetLastError());
  return finalize_slots(ctx, slot, 1, false);
}

int reduce_norm_l8(gh_ctx* ctx, int prec, int64_t size, const void* x, int slot) {
  double* part = ctx->partials + (int64_t)slot * kRedBlocks;
  if (prec == GH_F64)
    hipLaunchKernelGGL(k_amax_part<double>, dim3(kRedBlocks), dim3(kBlock), 0, ctx->stream, size,
                       (const double*)x, part);
  else
    hipLaunchKernelGGL(k_amax_part<float>, dim3(kRedBlocks), dim3(kBlock), 0, ctx->stream, size,
                       (const float*)x, part);
  GH_HIP(hipGetLastError());
  return finalize_slots(ctx, slot, 1, true);
}

// Modified Gram-Schmidt sub-step of GMRES::orthogonalization (gmres.cc:109-120):
//   d = fold of the previous step's partials (the same fold k_finalize does, so
//   the host's Hessenberg entry and the device's d are the same number),
//   vv += (-d) * prev, then the partials of vv.next (or vv.vv when next == nullptr).
__global__ __launch_bounds__(kBlock) void k_mgs_part(int64_t n, double* __restrict__ vv,
                                                     const double* __restrict__ prev,
                                                     const double* __restrict__ prev_part,
                                                     const double* __restrict__ next,
                                                     double* __restrict__ part) {
  __shared__ double dsh;
  double md = 0.0;
  if (prev) {
    const double d = sum_partials<false>(prev_part);
    if (threadIdx.x == 0) dsh = d;
    __syncthreads();
    md = -dsh;
  }
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (; i + 3 * stride < n; i += 4 * stride) {
    double v[4], w[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) v[u] = vv[i + u * stride];
    if (prev) {
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        v[u] += md * prev[i + u * stride];
        vv[i + u * stride] = v[u];
      }
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) w[u] = next ? next[i + u * stride] : v[u];
#pragma unroll
    for (int u = 0; u < 4; ++u) acc[u] += v[u] * w[u];
  }
  for (; i < n; i += stride) {
    double v = vv[i];
    if (prev) {
      v += md * prev[i];
      vv[i] = v;
    }
    acc[0] += v * (next ? next[i] : v);
  }
  const double a = block_reduce<false>((acc[0] + acc[1]) + (acc[2] + acc[3]));
  if (threadIdx.x == 0) part[blockIdx.x] = a;
}

// Classical Gram-Schmidt pieces of the flexible GMRES Arnoldi step: one pass
// forms the partials of <w, basis[k]> for up to kCgsGroup basis vectors (w
// read once), a second pass subtracts sum_k h_k basis[k] from w (k ascending)
// and forms the partials of <w, w>.  Per Arnoldi step that is two passes over
// w against MGS's one per basis vector.
constexpr int kCgsGroup = 8;

__global__ __launch_bounds__(kBlock) void k_cgs_dot(int64_t n, const double* __restrict__ w,
                                                    const double* const* __restrict__ basis,
                                                    int k0, int cnt, double* __restrict__ part) {
  const double* v[kCgsGroup];
#pragma unroll
  for (int g = 0; g < kCgsGroup; ++g) v[g] = g < cnt ? basis[k0 + g] : nullptr;
  double acc[kCgsGroup];
#pragma unroll
  for (int g = 0; g < kCgsGroup; ++g) acc[g] = 0.0;
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) {
    const double wi = w[i];
#pragma unroll
    for (int g = 0; g < kCgsGroup; ++g)
      if (g < cnt) acc[g] += wi * v[g][i];
  }
#pragma unroll
  for (int g = 0; g < kCgsGroup; ++g) {
    if (g >= cnt) break;  // cnt is uniform: every thread leaves at the same g
    const double a = block_reduce<false>(acc[g]);
    if (threadIdx.x == 0) part[(int64_t)(k0 + g) * kRedBlocks + blockIdx.x] = a;
  }
}

__global__ __launch_bounds__(kBlock) void k_cgs_update(int64_t n, double* __restrict__ w,
                                                       const double* const* __restrict__ basis,
                                                       int dim, const double* __restrict__ h,
                                                       double* __restrict__ part) {
  __shared__ double hs[256];
  __shared__ const double* vs[256];
  for (int k = threadIdx.x; k < dim; k += kBlock) {
    hs[k] = h[k];
    vs[k] = basis[k];
  }
  __syncthreads();
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  double acc = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) {
    double x = w[i];
    for (int k = 0; k < dim; ++k) x -= hs[k] * vs[k][i];
    w[i] = x;
    acc += x * x;
  }
  const double a = block_reduce<false>(acc);
  if (threadIdx.x == 0) part[blockIdx.x] = a;
}

int cgs_dots(gh_ctx* ctx, int64_t size, const double* w, const double* const* dbasis, int dim,
             int slot0) {
  for (int k0 = 0; k0 < dim; k0 += kCgsGroup)
    hipLaunchKernelGGL(k_cgs_dot, dim3(kRedBlocks), dim3(kBlock), 0, ctx->stream, size, w, dbasis,
                       k0, std::min(kCgsGroup, dim - k0), ctx->partials + (int64_t)slot0 * kRedBlocks);
  GH_HIP(hipGetLastError());
  return finalize_slots(ctx, slot0, dim, false);
}

int cgs_update(gh_ctx* ctx, int64_t size, double* w, const double* const* dbasis, int dim,
               int hslot0, int oslot) {
  if (dim > 256) return fail(GH_EINVAL, "cgs_update: at most 256 basis vectors");
  hipLaunchKernelGGL(k_cgs_update, dim3(kRedBlocks), dim3(kBlock), 0, ctx->stream, size, w, dbasis,
                     dim, ctx->scal + hslot0, ctx->partials + (int64_t)oslot * kRedBlocks);
  GH_HIP(hipGetLastError());
  return finalize_slots(ctx, oslot, 1, false);
}

int mgs_step(gh_ctx* ctx, int64_t size, double* vv, const double* prev, int dslot,
             const double* next, int oslot) {
  hipLaunchKernelGGL(k_mgs_part, dim3(kRedBlocks), dim3(kBlock), 0, ctx->stream, size, vv, prev,
                     prev ? ctx->partials + (int64_t)dslot * kRedBlocks : nullptr, next,
                     ctx->partials + (int64_t)oslot * kRedBlocks);
  GH_HIP(hipGetLastError());
  return GH_OK;
}

int vec_axpby(gh_ctx* ctx, int prec, int64_t size, void* y, double a, double d, const void* x) {
  // mode from (a, x): a==1 -> add; a==0 && x -> equ
  int mode = (a == 1.0) ? 1 : 0;
  if (prec == GH_F64)
    hipLaunchKernelGGL(k_axpby<double>, dim3(grid_stride(size)), dim3(kBlock), 0, ctx->stream,
                       size, (double*)y, a, d, (const double*)x, mode);
  else
    hipLaunchKernelGGL(k_axpby<float>, dim3(grid_stride(size)), dim3(kBlock), 0, ctx->stream,
                       size, (float*)y, (float)a, (float)d, (const float*)x, mode);
  GH_HIP(hipGetLastError());
  return GH_OK;
}

int vec_scale_to(gh_ctx* ctx, int prec, int64_t size, void* y, double s, const void* x) {
  if (prec == GH_F64)
    hipLaunchKernelGGL(k_axpby<double>, dim3(grid_stride(size)), dim3(kBlock), 0, ctx->stream,
                       size, (double*)y, 0.0, s, (const double*)x, 2);
  else
    hipLaunchKernelGGL(k_axpby<float>, dim3(grid_stride(size)), dim3(kBlock), 0, ctx->stream,
                       size, (float*)y, 0.f, (float)s, (const float*)x, 2);
  GH_HIP(hipGetLastError());
  return GH_OK;
}

int vec_convert(gh_ctx* ctx, int64_t size, int dprec, void* dst, int sprec, const void* src) {
  unsigned g = grid_stride(size);
  if (dprec == GH_F64 && sprec == GH_F32)
    hipLaunchKernelGGL((k_convert<double, float>), dim3(g), dim3(kBlock), 0, ctx->stream, size,
                       (double*)dst, (const float*)src);
  else if (dprec == GH_F32 && sprec == GH_F64)
    hipLaunchKernelGGL((k_convert<float, double>), dim3(g), dim3(kBlock), 0, ctx->stream, size,
                       (float*)dst, (const double*)src);
  else if (dprec == sprec) {
    // a kernel rather than a memcpy node, so that the copy can sit in a captured graph
    if (dst == src) return GH_OK;
    if (dprec == GH_F64)
      hipLaunchKernelGGL((k_convert<double, double>), dim3(g), dim3(kBlock), 0, ctx->stream,
                         size, (double*)dst, (const double*)src);
    else
      hipLaunchKernelGGL((k_convert<float, float>), dim3(g), dim3(kBlock), 0, ctx->stream, size,
                         (float*)dst, (const float*)src);
  } else
    return fail(GH_EINVAL, "vec_convert: unsupported precisions %d <- %d", dprec, sprec);
  GH_HIP(hipGetLastError());
  return GH_OK;
}

int vec_fill_zero(gh_ctx* ctx, int prec, int64_t size, void* x) {
  if (prec == GH_F64)
    hipLaunchKernelGGL(k_fill0<double>, dim3(grid_stride(size)), dim3(kBlock), 0, ctx->stream,
                       size, (double*)x);
  else
    hipLaunchKernelGGL(k_fill0<float>, dim3(grid_stride(size)), dim3(kBlock), 0, ctx->stream,
                       size, (float*)x);
  GH_HIP(hipGetLastError());
  return GH_OK;
}

// ------------------------------------------------------------------ SpMV
// block matrices (ncomp > 1): one thread per block row component
template <int B, typename XT, typename Acc, class Op>
static void launch_bsell(gh_ctx* ctx, const BSell& S, const XT* x, Op op) {
  const unsigned grid = grid_for(S.nslices * 64);
  if (S.prec == GH_F64)
    hipLaunchKernelGGL((k_bsell<B, double, XT, Acc, Op>), dim3(grid), dim3(kBlock), 0, ctx->stream,
                       S.nslices, S.sptr, S.row, S.rlen, S.col, (const double*)S.val, x, op);
  else if (S.prec == GH_F32)
    hipLaunchKernelGGL((k_bsell<B, float, XT, Acc, Op>), dim3(grid), dim3(kBlock), 0, ctx->stream,
                       S.nslices, S.sptr, S.row, S.rlen, S.col, (const float*)S.val, x, op);
  else
    hipLaunchKernelGGL((k_bsell<B, __half, XT, Acc, Op>), dim3(grid), dim3(kBlock), 0, ctx->stream,
                       S.nslices, S.sptr, S.row, S.rlen, S.col, (const __half*)S.val, x, op);
}

template <typename XT, typename Acc, class Op>
static int dispatch_val(gh_ctx* ctx, const gh_csr* A, const XT* x, Op op) {
  // component-row kernel where it is built and x is aligned for its block loads
  const BSell& S = A->bsell;
  if (S.B == A->ncomp && ((uintptr_t)x % (sizeof(XT) * (S.B == 4 ? 4 : 1))) == 0) {
    if (S.nslices == 0) return GH_OK;
    if (S.B == 4) launch_bsell<4, XT, Acc>(ctx, S, x, op);
    else if (S.B == 3) launch_bsell<3, XT, Acc>(ctx, S, x, op);
    else launch_bsell<2, XT, Acc>(ctx, S, x, op);
    GH_HIP(hipGetLastError());
    return GH_OK;
  }
  unsigned grid = grid_for(A->n * A->ncomp);
  if (A->prec == GH_F64)
    hipLaunchKernelGGL((k_bsr<double, XT, Acc, Op>), dim3(grid), dim3(kBlock), 0, ctx->stream,
                       A->n, A->ncomp, A->rowptr, A->col, (const double*)A->val, x, op);
  else if (A->prec == GH_F32)
    hipLaunchKernelGGL((k_bsr<float, XT, Acc, Op>), dim3(grid), dim3(kBlock), 0, ctx->stream,
                       A->n, A->ncomp, A->rowptr, A->col, (const float*)A->val, x, op);
  else
    hipLaunchKernelGGL((k_bsr<__half, XT, Acc, Op>), dim3(grid), dim3(kBlock), 0, ctx->stream,
                       A->n, A->ncomp, A->rowptr, A->col, (const __half*)A->val, x, op);
  GH_HIP(hipGetLastError());
  return GH_OK;
}

int csr_spmv(gh_ctx* ctx, const gh_csr* A, int yprec, void* y, int xprec, const void* x,
             double d, double b) {
  const bool ry = (b != 0.0);
  if (A->ncomp == 1) {
    // FP64 arithmetic in the stored row order, uncontracted: the reference's sums
    if (xprec == GH_F64 && yprec == GH_F64)
      return sell_dispatch<double, double>(ctx, A->sell, (const double*)x,
                                           SpmvOp<double, double, true>{(double*)y, d, b, ry});
    if (xprec == GH_F64 && yprec == GH_F32)
      return sell_dispatch<double, double>(ctx, A->sell, (const double*)x,
                                           SpmvOp<float, double, true>{(float*)y, d, b, ry});
    if (xprec == GH_F32 && yprec == GH_F32)
      return sell_dispatch<float, float>(ctx, A->sell, (const float*)x,
                                         SpmvOp<float, float, false>{(float*)y, (float)d, (float)b, ry});
    if (xprec == GH_F32 && yprec == GH_F64)
      return sell_dispatch<float, float>(ctx, A->sell, (const float*)x,
                                         SpmvOp<double, float, false>{(double*)y, (float)d, (float)b, ry});
  }
  if (xprec == GH_F64 && yprec == GH_F64)
    return dispatch_val<double, double>(ctx, A, (const double*)x,
                                        BsrSpmvOp<double, double>{(double*)y, d, b, ry});
  if (xprec == GH_F64 && yprec == GH_F32)
    return dispatch_val<double, double>(ctx, A, (const double*)x,
                                        BsrSpmvOp<float, double>{(float*)y, d, b, ry});
  if (xprec == GH_F32 && yprec == GH_F32)
    return dispatch_val<float, float>(ctx, A, (const float*)x,
                                      BsrSpmvOp<float, float>{(float*)y, (float)d, (float)b, ry});
  if (xprec == GH_F32 && yprec == GH_F64)
    return dispatch_val<float, float>(ctx, A, (const float*)x,
                                      BsrSpmvOp<double, float>{(double*)y, (float)d, (float)b, ry});
  return fail(GH_EINVAL, "csr_spmv: bad precisions");
}

// y = A x with FP32 x and FP64 values and arithmetic (flexible GMRES products)
int csr_spmv_x32(gh_ctx* ctx, const gh_csr* A, double* y, const float* x) {
  if (A->prec != GH_F64) return fail(GH_EINVAL, "csr_spmv_x32: needs an FP64 matrix");
  if (A->ncomp == 1)
    return sell_dispatch<float, double>(ctx, A->sell, x,
                                        SpmvOp<double, double, true>{y, 1.0, 0.0, false});
  return dispatch_val<float, double>(ctx, A, x, BsrSpmvOp<double, double>{y, 1.0, 0.0, false});
}

int csr_residual(gh_ctx* ctx, const gh_csr* A, int prec, void* r, const void* x, const void* b) {
  if (A->ncomp == 1) {
    if (prec == GH_F64)
      return sell_dispatch<double, double>(ctx, A->sell, (const double*)x,
                                           ResidualOp<double, true>{(double*)r, (const double*)b});
    return sell_dispatch<float, float>(ctx, A->sell, (const float*)x,
                                       ResidualOp<float, false>{(float*)r, (const float*)b});
  }
  if (prec == GH_F64)
    return dispatch_val<double, double>(ctx, A, (const double*)x,
                                        BsrResidualOp<double>{(double*)r, (const double*)b});
  return dispatch_val<float, float>(ctx, A, (const float*)x,
                                    BsrResidualOp<float>{(float*)r, (const float*)b});
}

template <typename T, typename DT>
static int jacobi_impl(gh_ctx* ctx, const gh_csr* A, T* xn, const T* x, const T* b, T omega,
                       int smoother, bool x_zero) {
  const DT* dinv = (const DT*)A->dinv;
  if (A->ncomp > 1) {
    if (smoother != 0) return fail(GH_EINVAL, "block matrices support the jacobi smoother only");
    unsigned grid = grid_for(A->n);
    if (A->prec == GH_F64)
      hipLaunchKernelGGL((k_bsr_jacobi<double, T, DT>), dim3(grid), dim3(kBlock), 0, ctx->stream,
                         A->n, A->ncomp, A->rowptr, A->col, (const double*)A->val, x, b, dinv,
                         xn, omega, (int)x_zero);
    else if (A->prec == GH_F32)
      hipLaunchKernelGGL((k_bsr_jacobi<float, T, DT>), dim3(grid), dim3(kBlock), 0, ctx->stream,
                         A->n, A->ncomp, A->rowptr, A->col, (const float*)A->val, x, b, dinv,
                         xn, omega, (int)x_zero);
    else
      hipLaunchKernelGGL((k_bsr_jacobi<__half, T, DT>), dim3(grid), dim3(kBlock), 0,
                         ctx->stream, A->n, A->ncomp, A->rowptr, A->col, (const __half*)A->val,
                         x, b, dinv, xn, omega, (int)x_zero);
    GH_HIP(hipGetLastError());
    return GH_OK;
  }
  if (x_zero) {
    hipLaunchKernelGGL((k_jacobi_zero<T, DT>), dim3(grid_stride(A->n)), dim3(kBlock), 0,
                       ctx->stream, A->n, xn, b, dinv, omega, smoother);
    GH_HIP(hipGetLastError());
    return GH_OK;
  }
  constexpr bool exact = std::is_same<T, double>::value;
  return sell_dispatch<T, T>(ctx, A->sell, x,
                             JacobiOp<T, DT, exact>{xn, x, b, dinv, omega, smoother});
}

int csr_jacobi_sweep(gh_ctx* ctx, const gh_csr* A, int prec, void* xnew, const void* x,
                     const void* b, double omega, int smoother, bool x_zero) {
  if (prec == GH_F64) {
    if (A->prec != GH_F64) return fail(GH_EINVAL, "FP64 sweep needs an FP64 matrix");
    return jacobi_impl<double, double>(ctx, A, (double*)xnew, (const double*)x,
                                       (const double*)b, omega, smoother, x_zero);
  }
  // the inverse diagonal is kept in FP64 for FP64 matrices, in FP32 otherwise
  if (A->prec == GH_F64) return fail(GH_EINVAL, "FP32 sweep needs an FP32 or FP16 matrix");
  return jacobi_impl<float, float>(ctx, A, (float*)xnew, (const float*)x, (const float*)b,
                                   (float)omega, smoother, x_zero);
}

// ------------------------------------------------------------- transfers
int transfer_restrict(gh_ctx* ctx, const gh_transfer* T, int ncomp, int prec, void* uc,
                      const void* uf, const int32_t* /*unused*/) {
  if (ncomp == 1) {
    if (prec == GH_F64)
      return sell_w_launch<double>(ctx, T->r_sell, (const double*)uf, RestrictOp<double>{(double*)uc});
    return sell_w_launch<float>(ctx, T->r_sell, (const float*)uf, RestrictOp<float>{(float*)uc});
  }
  unsigned grid = grid_for(T->ncoarse * ncomp);
  if (prec == GH_F64)
    hipLaunchKernelGGL(k_restrict<double>, dim3(grid), dim3(kBlock), 0, ctx->stream, T->ncoarse,
                       ncomp, T->r_row, T->r_col, (const double*)uf, (double*)uc);
  else
    hipLaunchKernelGGL(k_restrict<float>, dim3(grid), dim3(kBlock), 0, ctx->stream, T->ncoarse,
                       ncomp, T->r_row, T->r_col, (const float*)uf, (float*)uc);
  GH_HIP(hipGetLastError());
  return GH_OK;
}

int transfer_prolongate_add(gh_ctx* ctx, const gh_transfer* T, int ncomp, int prec, void* uf,
                            const void* uc) {
  if (ncomp == 1) {
    if (prec == GH_F64)
      return sell_w_launch<double>(ctx, T->p_sell, (const double*)uc,
                                   ProlongAddOp<double>{(double*)uf});
    return sell_w_launch<float>(ctx, T->p_sell, (const float*)uc, ProlongAddOp<float>{(float*)uf});
  }
  unsigned grid = grid_for(T->nfine * ncomp);
  if (prec == GH_F64)
    hipLaunchKernelGGL(k_prolongate_add<double>, dim3(grid), dim3(kBlock), 0, ctx->stream,
                       T->nfine, ncomp, T->p_row, T->p_col, (const double*)uc, (double*)uf);
  else
    hipLaunchKernelGGL(k_prolongate_add<float>, dim3(grid), dim3(kBlock), 0, ctx->stream,
                       T->nfine, ncomp, T->p_row, T->p_col, (const float*)uc, (float*)uf);
  GH_HIP(hipGetLastError());
  return GH_OK;
}

// op 0 average, 1 distribute, 2 zero
int hn_apply(gh_ctx* ctx, const gh_hn* H, int ncomp, int prec, void* x, int op) {
  if (H->ne + H->nf == 0) return GH_OK;
  if (op == 0) {
    unsigned grid = grid_for((H->ne + H->nf) * ncomp);
    if (prec == GH_F64)
      hipLaunchKernelGGL(k_hn_average<double>, dim3(grid), dim3(kBlock), 0, ctx->stream, H->kind,
                         H->ne, H->edges, H->nf, H->faces, ncomp, (double*)x);
    else
      hipLaunchKernelGGL(k_hn_average<float>, dim3(grid), dim3(kBlock), 0, ctx->stream, H->kind,
                         H->ne, H->edges, H->nf, H->faces, ncomp, (float*)x);
  } else {
    if (op == 1 && H->np > 0) {
      unsigned grid = grid_for(H->np * ncomp);
      if (prec == GH_F64)
        hipLaunchKernelGGL(k_hn_gather<double>, dim3(grid), dim3(kBlock), 0, ctx->stream, H->np,
                           ncomp, H->d_parent, H->d_row, H->d_src, H->d_w, (double*)x);
      else
        hipLaunchKernelGGL(k_hn_gather<float>, dim3(grid), dim3(kBlock), 0, ctx->stream, H->np,
                           ncomp, H->d_parent, H->d_row, H->d_src, H->d_w, (float*)x);
      GH_HIP(hipGetLastError());
    }
    unsigned grid = grid_for((H->ne + H->nf) * ncomp);
    if (prec == GH_F64)
      hipLaunchKernelGGL(k_zero_nodes<double>, dim3(grid), dim3(kBlock), 0, ctx->stream,
                         H->ne + H->nf, ncomp, H->hanging, (double*)x);
    else
      hipLaunchKernelGGL(k_zero_nodes<float>, dim3(grid), dim3(kBlock), 0, ctx->stream,
                         H->ne + H->nf, ncomp, H->hanging, (float*)x);
  }
  GH_HIP(hipGetLastError());
  return GH_OK;
}

int mask_apply(gh_ctx* ctx, const gh_mask* M, int prec, void* x) {
  const int64_t m = M->nidx * M->ncomps;
  if (m == 0) return GH_OK;
  if (prec == GH_F64)
    hipLaunchKernelGGL(k_zero_flat<double>, dim3(grid_for(m)), dim3(kBlock), 0, ctx->stream, m,
                       M->idx, (double*)x);
  else
    hipLaunchKernelGGL(k_zero_flat<float>, dim3(grid_for(m)), dim3(kBlock), 0, ctx->stream, m,
                       M->idx, (float*)x);
  GH_HIP(hipGetLastError());
  return GH_OK;
}

template <typename T>
static int upload(gh_ctx* ctx, T** dptr, const T* host, size_t count) {
  *dptr = nullptr;
  size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
  if (hipMalloc((void**)dptr, bytes) != hipSuccess)
    return fail(GH_ENOMEM, "hipMalloc(%zu) failed", bytes);
  if (count) GH_HIP(hipMemcpyAsync(*dptr, host, count * sizeof(T), hipMemcpyHostToDevice, ctx->stream));
  return GH_OK;
}

// A relative column pattern shared by at least this many rows goes into the
// pattern table; a slice runs from the table when all its rows do and they
// follow at most kMaxSlicePatterns different patterns (the per-lane offset
// loads then touch a few table lines per step).
constexpr int64_t kMinPatternRows = 16;
constexpr int kMaxSlicePatterns = 16;

// orig: original row of each slot (nullptr: slot = row); rowptr/col/vals are
// already in slot order.
// A table from 64-bit keys (the row-pattern hashes) to counters or row
// indices: open addressing with linear probing, sized for n keys; it stands
// in for std::unordered_map where millions of rows are looked up.
struct KeyTable {
  std::vector<uint64_t> key;
  std::vector<int64_t> val;
  std::vector<uint8_t> used;
  uint64_t mask = 0;
  KeyTable(int64_t n, int64_t init) {
    uint64_t cap = 16;
    while (cap < 2 * (uint64_t)std::max<int64_t>(n, 1)) cap <<= 1;
    key.assign(cap, 0);
    val.assign(cap, init);
    used.assign(cap, 0);
    mask = cap - 1;
    fill = init;
  }
  // the value of k, inserted with the initial value if absent
  int64_t& operator[](uint64_t k) {
    uint64_t h = (k * 0x9E3779B97F4A7C15ull) >> 17;
    for (;; ++h) {
      const uint64_t b = h & mask;
      if (!used[b]) {
        used[b] = 1;
        key[b] = k;
        val[b] = fill;
        return val[b];
      }
      if (key[b] == k) return val[b];
    }
  }
  int64_t fill = 0;
};

static int sell_build_impl(gh_ctx* ctx, int64_t n, const int32_t* rowptr, const uint32_t* col,
                           const double* vals, int prec, const int32_t* orig, Sell* S) {
  S->n = n;
  S->prec = prec;
  S->maxlen = 0;
  S->nslices = (n + kSlice - 1) / kSlice;
  const int64_t ns = S->nslices;
  const bool valued = vals != nullptr;
  // entries kept per row: a valued operator drops its exact zeros
  auto keep = [&](int64_t q) { return !valued || vals[q] != 0.0; };
  // The loops over rows and slices run on host threads; each writes its own
  // rows or slices, and the passes that accumulate (offsets, counters, the
  // pattern candidates) run in order, so the layout does not depend on the
  // thread count.
  std::vector<int32_t> lens(n);
  parallel_rows(n, [&](int64_t a, int64_t b) {
    for (int64_t i = a; i < b; ++i) {
      int32_t len = 0;
      for (int64_t q = rowptr[i]; q < rowptr[i + 1]; ++q) len += keep(q);
      lens[i] = len;
    }
  });
  std::vector<uint32_t> sptr(ns + 1, 0);
  std::vector<uint16_t> rlen(n);
  int64_t stored = 0;
  for (int64_t i = 0; i < n; ++i) {
    if (lens[i] > 65535) return fail(GH_EINVAL, "row %lld has more than 65535 entries", (long long)i);
    rlen[i] = (uint16_t)lens[i];
    stored += lens[i];
    S->maxlen = std::max<int>(S->maxlen, lens[i]);
  }
  std::vector<int32_t>().swap(lens);
  S->stored = stored;
  uint64_t total = 0;
  for (int64_t s = 0; s < ns; ++s) {
    int w = 0;
    for (int64_t i = s * kSlice; i < std::min<int64_t>(n, (s + 1) * kSlice); ++i)
      w = std::max<int>(w, rlen[i]);
    sptr[s] = (uint32_t)total;
    total += (uint64_t)w * kSlice;
    if (total > 0xFFFFFFFFull) return fail(GH_EINVAL, "operator too large for 32-bit slot offsets");
  }
  sptr[ns] = (uint32_t)total;
  S->slots = (int64_t)total;
  // per slot: column (global) and value, rows' kept entries in order
  std::vector<uint32_t> scol(total, 0);
  std::vector<double> sval(valued ? total : 0, 0.0);
  parallel_for(ns, 512, [&](int64_t s0, int64_t s1) {
    for (int64_t s = s0; s < s1; ++s)
      for (int lane = 0; lane < kSlice; ++lane) {
        const int64_t i = s * kSlice + lane;
        if (i >= n) break;
        uint64_t slot = sptr[s] + lane;
        for (int64_t q = rowptr[i]; q < rowptr[i + 1]; ++q) {
          if (!keep(q)) continue;
          scol[slot] = col[q];
          if (valued) sval[slot] = vals[q];
          slot += kSlice;
        }
      }
  });
  GH_TRY(upload(ctx, &S->sptr, sptr.data(), sptr.size()));
  if (!valued) {
    GH_TRY(upload(ctx, &S->rlen, rlen.data(), rlen.size()));
    GH_TRY(upload(ctx, &S->col, scol.data(), scol.size()));
    GH_HIP(hipStreamSynchronize(ctx->stream));
    return GH_OK;
  }
  // Relative column patterns (col - row over a row's kept entries, in order)
  // that at least kMinPatternRows rows share.  A slice whose rows all follow
  // table patterns stores no columns: rlen holds the pattern ids.
  std::vector<int32_t> rpid(n, -1), ptab, pstart(1, 0);
  {
    constexpr size_t kMaxCandidates = 1 << 20;
    constexpr int64_t kMaxTable = 1 << 22;  // offsets in the table
    auto orow = [&](int64_t i) { return orig ? (int64_t)orig[i] : i; };
    std::vector<uint64_t> rh(n);
    parallel_rows(n, [&](int64_t a, int64_t b) {
      for (int64_t i = a; i < b; ++i) {
        uint64_t hsh = 1469598103934665603ull;
        const int64_t o = orow(i);
        for (int64_t q = rowptr[i]; q < rowptr[i + 1]; ++q)
          if (keep(q)) hsh = (hsh ^ (uint32_t)(int32_t)((int64_t)col[q] - o)) * 1099511628211ull;
        rh[i] = hsh;
      }
    });
    // equal relative patterns of rows i and r
    auto same_pattern = [&](int64_t i, int64_t r) {
      int64_t q = rowptr[i], t = rowptr[r];
      const int64_t qe = rowptr[i + 1], te = rowptr[r + 1], oi = orow(i), orr = orow(r);
      for (;;) {
        while (q < qe && !keep(q)) ++q;
        while (t < te && !keep(t)) ++t;
        if (q == qe || t == te) return q == qe && t == te;
        if ((int64_t)col[q] - oi != (int64_t)col[t] - orr) return false;
        ++q;
        ++t;
      }
    };
    // rows per hash: only patterns frequent enough become candidates; the
    // first row of each such hash represents it
    std::vector<int32_t> rrep(n, -1);
    {
      KeyTable hcount(n, 0), first(n, -1);
      for (int64_t i = 0; i < n; ++i) hcount[rh[i]]++;
      for (int64_t i = 0; i < n; ++i)
        if (hcount[rh[i]] >= kMinPatternRows) {
          int64_t& f = first[rh[i]];
          if (f < 0) f = i;
          rrep[i] = (int32_t)f;
        }
    }
    std::vector<uint8_t> as_rep(n, 0);
    parallel_rows(n, [&](int64_t a, int64_t b) {
      for (int64_t i = a; i < b; ++i)
        if (rrep[i] >= 0) as_rep[i] = rrep[i] == i || same_pattern(i, rrep[i]);
    });
    // candidates in order of first appearance (the row that first shows each
    // pattern); a row like its hash's first row takes that row's candidate,
    // others (hash collisions) search the hash's candidates
    std::unordered_map<uint64_t, std::vector<int32_t>> byhash;
    std::vector<int32_t> cand_row, rcand(n, -1), cand_of(n, -1);
    std::vector<int64_t> count;
    for (int64_t i = 0; i < n; ++i) {
      if (rrep[i] < 0) continue;
      int32_t id = -1;
      if (as_rep[i] && rrep[i] != i) {
        id = cand_of[rrep[i]];
      } else {
        std::vector<int32_t>& ids = byhash[rh[i]];
        for (int32_t c : ids)
          if (same_pattern(i, cand_row[c])) id = c;
        if (id < 0 && cand_row.size() < kMaxCandidates) {
          id = (int32_t)cand_row.size();
          cand_row.push_back((int32_t)i);
          count.push_back(0);
          ids.push_back(id);
        }
      }
      cand_of[i] = id;
      if (id >= 0) {
        count[id]++;
        rcand[i] = id;
      }
    }
    // table in order of first appearance: patterns used by >= kMinPatternRows rows
    std::vector<int32_t> pid_of(cand_row.size(), -1);
    for (size_t c = 0; c < cand_row.size(); ++c) {
      const int64_t r = cand_row[c];
      if (count[c] < kMinPatternRows || (int64_t)ptab.size() + rlen[r] > kMaxTable ||
          pstart.size() > 65535)
        continue;
      pid_of[c] = (int32_t)pstart.size() - 1;
      for (int64_t q = rowptr[r]; q < rowptr[r + 1]; ++q)
        if (keep(q)) ptab.push_back((int32_t)((int64_t)col[q] - orow(r)));
      pstart.push_back((int32_t)ptab.size());
    }
    parallel_rows(n, [&](int64_t a, int64_t b) {
      for (int64_t i = a; i < b; ++i)
        if (rcand[i] >= 0) rpid[i] = pid_of[rcand[i]];
    });
  }
  S->npat = (int64_t)pstart.size() - 1;
  S->ntab = (int64_t)ptab.size();
  // column encoding per slice: pattern; else 16-bit offsets when the slice's
  // columns span fewer than 2^16 indices; else 32-bit.  Per slice first (on
  // the threads), then the offsets in slice order.
  std::vector<uint32_t> cptr(ns), cbase(ns, 0), lo_s(ns, 0), hi_s(ns, 0);
  std::vector<int64_t> kept_s(ns, 0);
  std::vector<uint8_t> wide(ns, 0), pat(ns, 0), one(ns, 0);
  parallel_for(ns, 512, [&](int64_t s0, int64_t s1) {
    std::vector<int32_t> ids(kSlice);
    for (int64_t s = s0; s < s1; ++s) {
      bool all = true;
      int64_t kept = 0;
      int m = 0;
      for (int lane = 0; lane < kSlice; ++lane) {
        const int64_t i = s * kSlice + lane;
        if (i >= n) break;
        all = all && rpid[i] >= 0;
        kept += rlen[i];
        ids[m++] = rpid[i];
      }
      kept_s[s] = kept;
      if (all) {
        bool same = true;
        for (int l = 1; l < m; ++l) same = same && ids[l] == ids[0];
        std::sort(ids.begin(), ids.begin() + m);
        all = (std::unique(ids.begin(), ids.begin() + m) - ids.begin()) <= kMaxSlicePatterns;
        one[s] = same;
      }
      if (all) {
        pat[s] = 1;
        continue;
      }
      uint32_t lo = 0xFFFFFFFFu, hi = 0;
      for (int lane = 0; lane < kSlice; ++lane) {
        const int64_t i = s * kSlice + lane;
        if (i >= n) break;
        for (int k = 0; k < rlen[i]; ++k) {
          const uint32_t c = scol[sptr[s] + (uint64_t)k * kSlice + lane];
          lo = std::min(lo, c);
          hi = std::max(hi, c);
        }
      }
      if (lo > hi) lo = hi = 0;  // no entries
      lo_s[s] = lo;
      hi_s[s] = hi;
    }
  });
  uint64_t units = 0;
  S->narrow = 0;
  S->patterned = 0;
  S->col_bytes = 0;
  for (int64_t s = 0; s < ns; ++s) {
    if (pat[s]) {
      cptr[s] = 0x40000000u;
      S->patterned++;
      if (one[s]) {  // one pattern for the whole wavefront: its id in cbase
        cptr[s] |= 0x20000000u;
        cbase[s] = (uint32_t)rpid[s * kSlice];
      }
      continue;
    }
    const uint64_t w = (sptr[s + 1] - sptr[s]);
    wide[s] = (hi_s[s] - lo_s[s]) > 0xFFFFu;
    if (wide[s]) {
      units += units & 1;  // 4-byte alignment
      cptr[s] = (uint32_t)units | 0x80000000u;
      units += 2 * w;
      S->col_bytes += 4 * kept_s[s];
    } else {
      cptr[s] = (uint32_t)units;
      cbase[s] = lo_s[s];
      units += w;
      S->narrow++;
      S->col_bytes += 2 * kept_s[s];
    }
    if (units >= 0x20000000ull) return fail(GH_EINVAL, "operator too large for its column index");
  }
  std::vector<uint16_t> cidx(std::max<uint64_t>(units, 1), 0);
  parallel_for(ns, 512, [&](int64_t s0, int64_t s1) {
    for (int64_t s = s0; s < s1; ++s) {
      if (pat[s]) {  // rlen holds the pattern ids
        for (int lane = 0; lane < kSlice && s * kSlice + lane < n; ++lane)
          rlen[s * kSlice + lane] = (uint16_t)rpid[s * kSlice + lane];
        continue;
      }
      const uint64_t w = (sptr[s + 1] - sptr[s]);
      const uint64_t at = cptr[s] & 0x1FFFFFFFu;
      for (uint64_t j = 0; j < w; ++j) {
        const int lane = (int)(j % kSlice);
        const int64_t i = s * kSlice + lane;
        const bool used = i < n && (int64_t)(j / kSlice) < rlen[i];
        const uint32_t c = used ? scol[sptr[s] + j] : cbase[s];  // padding: never read
        if (wide[s]) {
          cidx[at + 2 * j] = (uint16_t)(c & 0xFFFFu);
          cidx[at + 2 * j + 1] = (uint16_t)(c >> 16);
        } else {
          cidx[at + j] = (uint16_t)(c - cbase[s]);
        }
      }
    }
  });
  GH_TRY(upload(ctx, &S->rlen, rlen.data(), rlen.size()));
  GH_TRY(upload(ctx, &S->cidx, cidx.data(), cidx.size()));
  GH_TRY(upload(ctx, &S->cptr, cptr.data(), cptr.size()));
  GH_TRY(upload(ctx, &S->cbase, cbase.data(), cbase.size()));
  GH_TRY(upload(ctx, &S->ptab, ptab.data(), ptab.size()));
  GH_TRY(upload(ctx, &S->pstart, pstart.data(), pstart.size()));
  if (prec == GH_F64) {
    double* d = nullptr;
    GH_TRY(upload(ctx, &d, sval.data(), sval.size()));
    S->val = d;
  } else {
    double* tmp = nullptr;
    GH_TRY(upload(ctx, &tmp, sval.data(), sval.size()));
    const size_t m = std::max<size_t>(sval.size(), 1);
    if (hipMalloc(&S->val, m * (prec == GH_F32 ? 4 : 2)) != hipSuccess) {
      (void)hipFree(tmp);
      return fail(GH_ENOMEM, "sell_build: allocation failed");
    }
    if (prec == GH_F32)
      hipLaunchKernelGGL((k_convert<float, double>), dim3(grid_stride(total)), dim3(kBlock), 0,
                         ctx->stream, (int64_t)total, (float*)S->val, (const double*)tmp);
    else
      hipLaunchKernelGGL((k_convert<__half, double>), dim3(grid_stride(total)), dim3(kBlock), 0,
                         ctx->stream, (int64_t)total, (__half*)S->val, (const double*)tmp);
    GH_HIP(hipStreamSynchronize(ctx->stream));
    (void)hipFree(tmp);
  }
  GH_HIP(hipStreamSynchronize(ctx->stream));
  return GH_OK;
}

// Valued operators: the rows of a window of kSellWindow rows may run in
// another order (a stable sort), so that a slice holds rows of one length
// (no lanes idle on padding) and, where rows share a relative column pattern,
// one pattern (offsets from scalar loads).  Two candidate orders: longest rows
// first with equal lengths grouped by pattern, or longest rows first only; the
// first that clearly pays (5% fewer padded slots or 5% more single-pattern
// slices) and keeps each slice's rows close together is used, else the
// caller's order.  Each row keeps its entries and its sum; only the slot it
// runs in moves.
int sell_build(gh_ctx* ctx, int64_t n, const int32_t* rowptr, const uint32_t* col,
               const double* vals, int prec, Sell* S) {
  S->perm = nullptr;
  if (!vals || n < 2 * kSlice) return sell_build_impl(ctx, n, rowptr, col, vals, prec, nullptr, S);
  constexpr int64_t kSellWindow = 4096;
  std::vector<uint64_t> key(n);
  std::vector<int32_t> len(n);
  parallel_rows(n, [&](int64_t a, int64_t b) {
    for (int64_t i = a; i < b; ++i) {
      uint64_t h = 1469598103934665603ull;
      int32_t k = 0;
      for (int64_t q = rowptr[i]; q < rowptr[i + 1]; ++q) {
        if (vals[q] == 0.0) continue;
        h = (h ^ (uint32_t)(int32_t)((int64_t)col[q] - i)) * 1099511628211ull;
        ++k;
      }
      key[i] = h;
      len[i] = k;
    }
  });
  // rows of patterns frequent enough for the pattern table (by hash)
  std::vector<uint8_t> frequent(n, 0);
  {
    KeyTable hcount(n, 0);
    for (int64_t i = 0; i < n; ++i) hcount[key[i]]++;
    for (int64_t i = 0; i < n; ++i) frequent[i] = hcount[key[i]] >= kMinPatternRows;
  }
  // by_pattern: rows of one length with the frequent patterns first, then by
  // pattern, so that table rows fill slices of their own
  auto window_sort = [&](bool by_pattern) {
    std::vector<int32_t> p(n);
    for (int64_t i = 0; i < n; ++i) p[i] = (int32_t)i;
    const int64_t nw = (n + kSellWindow - 1) / kSellWindow;
    parallel_for(nw, 16, [&](int64_t wa, int64_t wb) {
      for (int64_t w = wa; w < wb; ++w) {
        const int64_t w0 = w * kSellWindow, w1 = std::min<int64_t>(n, w0 + kSellWindow);
        std::stable_sort(p.begin() + w0, p.begin() + w1, [&](int32_t a, int32_t b) {
          if (len[a] != len[b]) return len[a] > len[b];
          if (!by_pattern) return false;
          if (frequent[a] != frequent[b]) return frequent[a] > frequent[b];
          return key[a] < key[b];
        });
      }
    });
    return p;
  };
  // slices whose 64 rows share one pattern
  const int64_t nsl = (n + kSlice - 1) / kSlice;
  // per-slice figures on the threads, summed in slice order
  auto per_slice = [&](auto&& fn) {
    std::vector<int64_t> v(nsl);
    parallel_for(nsl, 1024, [&](int64_t a, int64_t b) {
      for (int64_t s = a; s < b; ++s) v[s] = fn(s * kSlice);
    });
    int64_t t = 0;
    for (int64_t x : v) t += x;
    return t;
  };
  auto uniform = [&](const std::vector<int32_t>* p) {
    return per_slice([&](int64_t s0) -> int64_t {
      if (s0 + kSlice > n) return 0;
      const uint64_t k0 = key[p ? (*p)[s0] : s0];
      bool same = true;
      for (int l = 1; l < kSlice && same; ++l) same = key[p ? (*p)[s0 + l] : s0 + l] == k0;
      return same;
    });
  };
  // slots a sweep runs: 64 times the longest row of each slice
  auto slots = [&](const std::vector<int32_t>* p) {
    return per_slice([&](int64_t s0) -> int64_t {
      int32_t w = 0;
      for (int64_t t0 = s0; t0 < std::min<int64_t>(n, s0 + kSlice); ++t0)
        w = std::max(w, len[p ? (*p)[t0] : t0]);
      return (int64_t)kSlice * w;
    });
  };
  // rows a slice's lanes touch, in groups of 16 (the vectors' 128-byte lines
  // in FP64): 4 or 5 for 64 consecutive rows
  auto groups = [&](const std::vector<int32_t>& p) {
    const int64_t g = per_slice([&](int64_t s0) -> int64_t {
      int32_t b[kSlice];
      const int m = (int)std::min<int64_t>(kSlice, n - s0);
      for (int l = 0; l < m; ++l) b[l] = p[s0 + l] >> 4;
      std::sort(b, b + m);
      return std::unique(b, b + m) - b;
    });
    return (double)g / (double)nsl;
  };
  const int64_t u0 = uniform(nullptr), t0 = slots(nullptr);
  const int64_t du = std::max<int64_t>(1, (n / kSlice) / 20);
  // worth it only when clearly fewer slots run or clearly more slices hold one
  // pattern, with rows that stay close together (rows of two kinds interleaved
  // at stride 2 touch ~8-16 groups per slice and gain; rows of many kinds
  // grouped from all over the window touch 20+ and lose the vectors'
  // coalescing)
  auto pays = [&](const std::vector<int32_t>& p, bool count_patterns) {
    const bool fewer = slots(&p) * 20 <= t0 * 19;
    const bool more = count_patterns && uniform(&p) >= u0 + du;
    return (fewer || more) && groups(p) <= 20.0;
  };
  std::vector<int32_t> perm = window_sort(true);
  if (!pays(perm, true)) {
    perm = window_sort(false);
    if (!pays(perm, false)) return sell_build_impl(ctx, n, rowptr, col, vals, prec, nullptr, S);
  }
  // the operator in slot order
  std::vector<int32_t> rp2(n + 1, 0);
  for (int64_t t = 0; t < n; ++t) rp2[t + 1] = rp2[t] + (rowptr[perm[t] + 1] - rowptr[perm[t]]);
  std::vector<uint32_t> c2(rp2[n]);
  std::vector<double> v2(rp2[n]);
  parallel_rows(n, [&](int64_t a, int64_t b) {
    for (int64_t t = a; t < b; ++t) {
      const int32_t i = perm[t];
      std::copy(col + rowptr[i], col + rowptr[i + 1], c2.begin() + rp2[t]);
      std::copy(vals + rowptr[i], vals + rowptr[i + 1], v2.begin() + rp2[t]);
    }
  });
  GH_TRY(sell_build_impl(ctx, n, rp2.data(), c2.data(), v2.data(), prec, perm.data(), S));
  GH_TRY(upload(ctx, &S->perm, perm.data(), perm.size()));
  GH_HIP(hipStreamSynchronize(ctx->stream));
  return GH_OK;
}

void sell_free(Sell* S) {
  for (void* p : {(void*)S->sptr, (void*)S->rlen, (void*)S->col, S->val, (void*)S->cidx,
                  (void*)S->cptr, (void*)S->cbase, (void*)S->ptab, (void*)S->pstart,
                  (void*)S->perm})
    if (p) (void)hipFree(p);
  S->perm = nullptr;
  S->ptab = nullptr;
  S->pstart = nullptr;
  S->sptr = nullptr;
  S->rlen = nullptr;
  S->col = nullptr;
  S->cidx = nullptr;
  S->cptr = nullptr;
  S->cbase = nullptr;
  S->val = nullptr;
}

int bsell_build(gh_ctx* ctx, int64_t n, int B, const int32_t* rowptr, const int32_t* col,
                const double* vals, int prec, BSell* S) {
  const int BB = B * B;
  const int64_t nr = n * B;
  S->B = B;
  S->nrows = nr;
  S->prec = prec;
  auto kept = [&](int64_t q, int r) {
    const double* m = vals + q * BB + (int64_t)r * B;
    for (int c = 0; c < B; ++c)
      if (m[c] != 0.0) return true;
    return false;
  };
  std::vector<uint16_t> len(nr);
  int64_t stored = 0;
  for (int64_t i = 0; i < n; ++i)
    for (int r = 0; r < B; ++r) {
      int64_t c = 0;
      for (int32_t q = rowptr[i]; q < rowptr[i + 1]; ++q) c += kept(q, r);
      if (c > 65535) return fail(GH_EINVAL, "block row %lld has more than 65535 entries", (long long)i);
      len[i * B + r] = (uint16_t)c;
      stored += c;
    }
  // longest rows first inside each window, so that slices hold rows of equal length
  std::vector<int32_t> order(nr);
  for (int64_t t = 0; t < nr; ++t) order[t] = (int32_t)t;
  for (int64_t w0 = 0; w0 < nr; w0 += kBSellWindow) {
    const int64_t w1 = std::min<int64_t>(nr, w0 + kBSellWindow);
    std::stable_sort(order.begin() + w0, order.begin() + w1,
                     [&](int32_t a, int32_t b) { return len[a] > len[b]; });
  }
  S->nslices = (nr + 63) / 64;
  std::vector<uint32_t> sptr(S->nslices + 1, 0);
  std::vector<int32_t> lrow(S->nslices * 64, -1);
  std::vector<uint16_t> llen(S->nslices * 64, 0);
  uint64_t units = 0;
  for (int64_t s = 0; s < S->nslices; ++s) {
    int w = 0;
    for (int l = 0; l < 64 && s * 64 + l < nr; ++l) {
      const int32_t t = order[s * 64 + l];
      lrow[s * 64 + l] = t;
      llen[s * 64 + l] = len[t];
      w = std::max<int>(w, len[t]);
    }
    sptr[s] = (uint32_t)units;
    units += (uint64_t)w;
    if (units > 0xFFFFFFFFull) return fail(GH_EINVAL, "block operator too large for its slot offsets");
  }
  sptr[S->nslices] = (uint32_t)units;
  S->slots = (int64_t)units * 64;
  S->stored = stored;
  std::vector<int32_t> scol(std::max<int64_t>(S->slots, 1), 0);
  std::vector<double> sval(std::max<int64_t>(S->slots * B, 1), 0.0);
  for (int64_t s = 0; s < S->nslices; ++s)
    for (int l = 0; l < 64; ++l) {
      const int32_t t = lrow[s * 64 + l];
      if (t < 0) continue;
      const int64_t i = t / B;
      const int r = (int)(t - i * B);
      int64_t k = 0;
      for (int32_t q = rowptr[i]; q < rowptr[i + 1]; ++q) {
        if (!kept(q, r)) continue;
        const int64_t slot = (int64_t)sptr[s] + k;
        scol[slot * 64 + l] = col[q];
        for (int c = 0; c < B; ++c)
          sval[(slot * B + c) * 64 + l] = vals[(int64_t)q * BB + (int64_t)r * B + c];
        ++k;
      }
    }
  GH_TRY(upload(ctx, &S->sptr, sptr.data(), sptr.size()));
  GH_TRY(upload(ctx, &S->row, lrow.data(), lrow.size()));
  GH_TRY(upload(ctx, &S->rlen, llen.data(), llen.size()));
  GH_TRY(upload(ctx, &S->col, scol.data(), scol.size()));
  double* tmp = nullptr;
  GH_TRY(upload(ctx, &tmp, sval.data(), sval.size()));
  if (prec == GH_F64) {
    S->val = tmp;
  } else {
    const size_t m = sval.size();
    if (hipMalloc(&S->val, m * (prec == GH_F32 ? 4 : 2)) != hipSuccess) {
      (void)hipFree(tmp);
      return fail(GH_ENOMEM, "bsell_build: allocation failed");
    }
    if (prec == GH_F32)
      hipLaunchKernelGGL((k_convert<float, double>), dim3(grid_stride((int64_t)m)), dim3(kBlock), 0,
                         ctx->stream, (int64_t)m, (float*)S->val, (const double*)tmp);
    else
      hipLaunchKernelGGL((k_convert<__half, double>), dim3(grid_stride((int64_t)m)), dim3(kBlock),
                         0, ctx->stream, (int64_t)m, (__half*)S->val, (const double*)tmp);
    GH_HIP(hipStreamSynchronize(ctx->stream));
    (void)hipFree(tmp);
  }
  GH_HIP(hipStreamSynchronize(ctx->stream));
  return GH_OK;
}

void bsell_free(BSell* S) {
  for (void* p : {(void*)S->sptr, (void*)S->row, (void*)S->rlen, (void*)S->col, S->val})
    if (p) (void)hipFree(p);
  S->sptr = nullptr;
  S->row = nullptr;
  S->rlen = nullptr;
  S->col = nullptr;
  S->val = nullptr;
  S->B = 0;
}

template <typename D, typename S>
__global__ __launch_bounds__(kBlock) void k_permute(int64_t n, int ncomp,
                                                    const int32_t* __restrict__ perm,
                                                    D* __restrict__ dst, const S* __restrict__ src,
                                                    int gather) {
  const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (t >= n * ncomp) return;
  const int64_t i = t / ncomp;
  const int c = (int)(t - i * ncomp);
  const int64_t o = (int64_t)perm[i] * ncomp + c;
  if (gather)
    dst[t] = (D)src[o];
  else
    dst[o] = (D)src[t];
}

int vec_permute(gh_ctx* ctx, int64_t n, int ncomp, const int32_t* perm, int dprec, void* dst,
                int sprec, const void* src, bool gather) {
  const unsigned g = grid_for(n * ncomp);
  const int gi = gather ? 1 : 0;
  if (dprec == GH_F64 && sprec == GH_F64)
    hipLaunchKernelGGL((k_permute<double, double>), dim3(g), dim3(kBlock), 0, ctx->stream, n, ncomp,
                       perm, (double*)dst, (const double*)src, gi);
  else if (dprec == GH_F32 && sprec == GH_F64)
    hipLaunchKernelGGL((k_permute<float, double>), dim3(g), dim3(kBlock), 0, ctx->stream, n, ncomp,
                       perm, (float*)dst, (const double*)src, gi);
  else if (dprec == GH_F64 && sprec == GH_F32)
    hipLaunchKernelGGL((k_permute<double, float>), dim3(g), dim3(kBlock), 0, ctx->stream, n, ncomp,
                       perm, (double*)dst, (const float*)src, gi);
  else
    hipLaunchKernelGGL((k_permute<float, float>), dim3(g), dim3(kBlock), 0, ctx->stream, n, ncomp,
                       perm, (float*)dst, (const float*)src, gi);
  GH_HIP(hipGetLastError());
  return GH_OK;
}

int transfer_from_csr(gh_ctx* ctx, int64_t ncoarse, int64_t nfine, std::vector<int32_t> rrow,
                      std::vector<uint32_t> rcol, std::vector<int32_t> prow,
                      std::vector<uint32_t> pcol, gh_transfer** out) {
  *out = nullptr;
  gh_transfer* T = new gh_transfer();
  T->ctx = ctx;
  T->ncoarse = ncoarse;
  T->nfine = nfine;
  T->h_rrow = std::move(rrow);
  T->h_rcol = std::move(rcol);
  T->h_prow = std::move(prow);
  T->h_pcol = std::move(pcol);
  int rc = upload(ctx, &T->r_row, T->h_rrow.data(), T->h_rrow.size());
  if (!rc) rc = upload(ctx, &T->r_col, T->h_rcol.data(), T->h_rcol.size());
  if (!rc) rc = upload(ctx, &T->p_row, T->h_prow.data(), T->h_prow.size());
  if (!rc) rc = upload(ctx, &T->p_col, T->h_pcol.data(), T->h_pcol.size());
  if (!rc) rc = sell_build(ctx, ncoarse, T->h_rrow.data(), T->h_rcol.data(), nullptr, GH_F64, &T->r_sell);
  if (!rc) rc = sell_build(ctx, nfine, T->h_prow.data(), T->h_pcol.data(), nullptr, GH_F64, &T->p_sell);
  ctx_retain(ctx);  // balanced by gh_transfer_destroy
  if (rc) {
    gh_transfer_destroy(T);
    return rc;
  }
  *out = T;
  return GH_OK;
}

}  // namespace gh

using namespace gh;

// ====================================================================== C ABI
extern "C" {

const char* gh_last_error(void) { return g_err.c_str(); }
const char* gh_version(void) { return "gascoigne_hip 0.1 (gfx950)"; }

int gh_ctx_create(int device, gh_ctx** out) {
  GH_ARG(out, "gh_ctx_create: out is NULL");
  *out = nullptr;
  int ndev = 0;
  GH_HIP(hipGetDeviceCount(&ndev));
  if (ndev < 1) return fail(GH_EHIP, "gh_ctx_create: no HIP device");
  if (device >= 0) {
    GH_ARG(device < ndev, "gh_ctx_create: device index out of range");
    GH_HIP(hipSetDevice(device));
  } else {
    GH_HIP(hipGetDevice(&device));
  }
  gh_ctx* c = new gh_ctx();
  c->device = device;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) c->num_cu = prop.multiProcessorCount;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
      hipMalloc((void**)&c->partials, sizeof(double) * kRedBlocks * kScalars) != hipSuccess ||
      hipMalloc((void**)&c->scal, sizeof(double) * kScalars) != hipSuccess ||
      hipHostMalloc((void**)&c->host_scal, sizeof(double) * kScalars, hipHostMallocDefault) !=
          hipSuccess) {
    ctx_release(c);
    return fail(GH_ENOMEM, "gh_ctx_create: allocation failed");
  }
  GH_HIP(hipMemsetAsync(c->scal, 0, sizeof(double) * kScalars, c->stream));
  GH_HIP(hipStreamSynchronize(c->stream));
  {
    // `device` is current here: look up its copy of the no-allocation
    // scratch once, and keep that address for every later context on it
    std::lock_guard<std::mutex> lock(g_noalloc_mu);
    if (device < kMaxDevices && !g_noalloc_ptr[device]) {
      void* p = nullptr;
      if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) == hipSuccess) g_noalloc_ptr[device] = p;
      else (void)hipGetLastError();  // gh_jacobi_noalloc then reports the failure
    }
    if (device < kMaxDevices) c->noalloc_scratch = g_noalloc_ptr[device];
  }
  *out = c;
  return GH_OK;
}

}  // extern "C"

namespace gh {
static void ctx_free(gh_ctx* ctx) {
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  direct_ctx_free(ctx);
  if (ctx->partials) (void)hipFree(ctx->partials);
  if (ctx->scal) (void)hipFree(ctx->scal);
  if (ctx->host_scal) (void)hipHostFree(ctx->host_scal);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}
void ctx_retain(gh_ctx* ctx) {
  if (ctx) ctx->refs.fetch_add(1);
}
void ctx_release(gh_ctx* ctx) {
  if (ctx && ctx->refs.fetch_sub(1) == 1) ctx_free(ctx);
}
}  // namespace gh

extern "C" {

int gh_ctx_destroy(gh_ctx* ctx) {
  ctx_release(ctx);
  return GH_OK;
}

int gh_ctx_sync(gh_ctx* ctx) {
  GH_ARG(ctx, "gh_ctx_sync: ctx is NULL");
  GH_HIP(hipStreamSynchronize(ctx->stream));
  return GH_OK;
}

void* gh_ctx_stream(gh_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

// ------------------------------------------------------------------ vectors
int gh_vec_create(gh_ctx* ctx, int64_t n, int ncomp, int prec, gh_vec** out) {
  GH_ARG(ctx && out, "gh_vec_create: NULL argument");
  gh_vec* v = new gh_vec();
  int rc = vec_alloc(ctx, n, ncomp, prec, v);
  if (rc) {
    delete v;
    return rc;
  }
  ctx_retain(ctx);
  *out = v;
  return GH_OK;
}

int gh_vec_destroy(gh_vec* v) {
  if (!v) return GH_OK;
  gh_ctx* ctx = v->ctx;
  if (ctx && ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  vec_free(v);
  delete v;
  ctx_release(ctx);
  return GH_OK;
}

int64_t gh_vec_n(const gh_vec* v) { return v ? v->n : -1; }
int gh_vec_ncomp(const gh_vec* v) { return v ? v->ncomp : -1; }
int gh_vec_prec(const gh_vec* v) { return v ? v->prec : -1; }
void* gh_vec_data(gh_vec* v) { return v ? v->d : nullptr; }

int gh_vec_upload(gh_ctx* ctx, gh_vec* v, const double* host) {
  GH_ARG(ctx && v && (host || v->size == 0), "gh_vec_upload: NULL argument");
  if (v->prec == GH_F64) {
    GH_HIP(hipMemcpyAsync(v->d, host, v->size * 8, hipMemcpyHostToDevice, ctx->stream));
    GH_HIP(hipStreamSynchronize(ctx->stream));
    return GH_OK;
  }
  // FP32 vector: rounded to nearest on the host (the same rounding as the
  // device conversion kernels), so that no device staging is needed
  std::vector<float> tmp((size_t)v->size);
  for (int64_t i = 0; i < v->size; ++i) tmp[i] = (float)host[i];
  GH_HIP(hipMemcpyAsync(v->d, tmp.data(), v->size * 4, hipMemcpyHostToDevice, ctx->stream));
  GH_HIP(hipStreamSynchronize(ctx->stream));
  return GH_OK;
}

int gh_vec_download(gh_ctx* ctx, const gh_vec* v, double* host) {
  GH_ARG(ctx && v && (host || v->size == 0), "gh_vec_download: NULL argument");
  if (v->prec == GH_F64) {
    GH_HIP(hipMemcpyAsync(host, v->d, v->size * 8, hipMemcpyDeviceToHost, ctx->stream));
    GH_HIP(hipStreamSynchronize(ctx->stream));
    return GH_OK;
  }
  // FP32 vector: copied as stored and widened (exactly) on the host
  std::vector<float> tmp((size_t)v->size);
  GH_HIP(hipMemcpyAsync(tmp.data(), v->d, v->size * 4, hipMemcpyDeviceToHost, ctx->stream));
  GH_HIP(hipStreamSynchronize(ctx->stream));
  for (int64_t i = 0; i < v->size; ++i) host[i] = (double)tmp[i];
  return GH_OK;
}

int gh_vec_zero(gh_ctx* ctx, gh_vec* v) {
  GH_ARG(ctx && v, "gh_vec_zero: NULL argument");
  return vec_fill_zero(ctx, v->prec, v->size, v->d);
}

static int same_shape(const gh_vec* a, const gh_vec* b) {
  return a && b && a->n == b->n && a->ncomp == b->ncomp;
}

int gh_vec_copy(gh_ctx* ctx, gh_vec* dst, const gh_vec* src) {
  GH_ARG(ctx && same_shape(dst, src), "gh_vec_copy: shape mismatch");
  return vec_convert(ctx, dst->size, dst->prec, dst->d, src->prec, src->d);
}

int gh_vec_equ(gh_ctx* ctx, gh_vec* dst, double s, const gh_vec* src) {
  GH_ARG(ctx && same_shape(dst, src) && dst->prec == src->prec, "gh_vec_equ: shape mismatch");
  return vec_scale_to(ctx, dst->prec, dst->size, dst->d, s, src->d);
}

int gh_vec_add(gh_ctx* ctx, gh_vec* dst, double s, const gh_vec* src) {
  GH_ARG(ctx && same_shape(dst, src) && dst->prec == src->prec, "gh_vec_add: shape mismatch");
  return vec_axpby(ctx, dst->prec, dst->size, dst->d, 1.0, s, src->d);
}

int gh_vec_sadd(gh_ctx* ctx, gh_vec* y, double a, double d, const gh_vec* v) {
  GH_ARG(ctx && same_shape(y, v) && y->prec == v->prec, "gh_vec_sadd: shape mismatch");
  if (y->prec == GH_F64)
    hipLaunchKernelGGL(k_axpby<double>, dim3(grid_stride(y->size)), dim3(kBlock), 0, ctx->stream,
                       y->size, (double*)y->d, a, d, (const double*)v->d, 0);
  else
    hipLaunchKernelGGL(k_axpby<float>, dim3(grid_stride(y->size)), dim3(kBlock), 0, ctx->stream,
                       y->size, (float*)y->d, (float)a, (float)d, (const float*)v->d, 0);
  GH_HIP(hipGetLastError());
  return GH_OK;
}

int gh_vec_scal(gh_ctx* ctx, gh_vec* x, double s) {
  GH_ARG(ctx && x, "gh_vec_scal: NULL argument");
  return vec_scale_to(ctx, x->prec, x->size, x->d, s, x->d);
}

int gh_vec_dot(gh_ctx* ctx, const gh_vec* x, const gh_vec* y, double* out) {
  GH_ARG(ctx && same_shape(x, y) && x->prec == y->prec && out, "gh_vec_dot: shape mismatch");
  GH_TRY(reduce_dot(ctx, x->prec, x->size, x->d, y->d, 0));
  return fetch_scalars(ctx, 0, 1, out);
}

int gh_vec_norm(gh_ctx* ctx, const gh_vec* x, double* out) {
  GH_ARG(ctx && x && out, "gh_vec_norm: NULL argument");
  GH_TRY(reduce_dot(ctx, x->prec, x->size, x->d, x->d, 0));
  double s;
  GH_TRY(fetch_scalars(ctx, 0, 1, &s));
  *out = std::sqrt(s);
  return GH_OK;
}

int gh_vec_norm_l8(gh_ctx* ctx, const gh_vec* x, double* out) {
  GH_ARG(ctx && x && out, "gh_vec_norm_l8: NULL argument");
  GH_TRY(reduce_norm_l8(ctx, x->prec, x->size, x->d, 0));
  return fetch_scalars(ctx, 0, 1, out);
}

}  // extern "C"

namespace gh {
// check_sorted = false admits rows whose entries are kept in a given order that
// is not ascending (operators renumbered for locality keep each row's
// original summation order)
template <typename DT>
static int launch_block_inverse(gh_ctx* ctx, int64_t n, int B, const int32_t* diagpos,
                                const double* val, DT* dinv, int* singular) {
  const dim3 g(grid_for(n)), blk(kBlock);
  switch (B) {
    case 1: hipLaunchKernelGGL((k_block_inverse<1, DT>), g, blk, 0, ctx->stream, n, diagpos, val, dinv, singular); break;
    case 2: hipLaunchKernelGGL((k_block_inverse<2, DT>), g, blk, 0, ctx->stream, n, diagpos, val, dinv, singular); break;
    case 3: hipLaunchKernelGGL((k_block_inverse<3, DT>), g, blk, 0, ctx->stream, n, diagpos, val, dinv, singular); break;
    case 4: hipLaunchKernelGGL((k_block_inverse<4, DT>), g, blk, 0, ctx->stream, n, diagpos, val, dinv, singular); break;
    case 5: hipLaunchKernelGGL((k_block_inverse<5, DT>), g, blk, 0, ctx->stream, n, diagpos, val, dinv, singular); break;
    case 6: hipLaunchKernelGGL((k_block_inverse<6, DT>), g, blk, 0, ctx->stream, n, diagpos, val, dinv, singular); break;
    case 7: hipLaunchKernelGGL((k_block_inverse<7, DT>), g, blk, 0, ctx->stream, n, diagpos, val, dinv, singular); break;
    case 8: hipLaunchKernelGGL((k_block_inverse<8, DT>), g, blk, 0, ctx->stream, n, diagpos, val, dinv, singular); break;
    default: return fail(GH_EINVAL, "block inverse: block size %d not in 1..8", B);
  }
  if (hipGetLastError() != hipSuccess) return fail(GH_EHIP, "block inverse launch");
  return GH_OK;
}

// ------------------------------------------------------- device assembly
// The SELL value stream of a scalar matrix from its FP64 CSR values: slot t
// (row perm[t]) stores the entries of its row that the layout kept, in CSR
// order, at sptr[s] + lane + 64 k; an entry left out at creation (an exact
// zero then) must still be zero, else flag[1] is set.  One wavefront per
// slice, in groups of 8 rows: the lanes read 64 consecutive CSR entries of a
// row at a time (coalesced; 4 rows in flight), rank the kept ones by a
// ballot and stage them in LDS, up to 128 per row (further ones are stored
// directly); the group then goes out k by k, the 8 rows of one k being 8
// adjacent slots.
template <typename VT>
__global__ __launch_bounds__(kBlock) void k_sell_refresh(int64_t n, const uint32_t* __restrict__ sptr,
                                                         const int32_t* __restrict__ perm,
                                                         const int32_t* __restrict__ rowptr,
                                                         const uint32_t* __restrict__ kept,
                                                         const double* __restrict__ val64,
                                                         VT* __restrict__ sval, int* flags) {
  constexpr int G = 8, KT = 128;
  __shared__ double tile[kBlock / kSlice][G][KT + 1];
  const int lane = threadIdx.x & (kSlice - 1);
  const int w = threadIdx.x >> 6;
  const int64_t s = (int64_t)blockIdx.x * (kBlock / kSlice) + w;
  const int64_t t0 = s * kSlice;
  if (t0 >= n) return;  // the whole wavefront
  const int rows = (int)min<int64_t>(kSlice, n - t0);
  // the lane's slot's row bounds, handed to the wavefront row by row below;
  // lanes past the last row hold empty ranges, so every row index below 64
  // is safe
  int32_t b0 = 0, b1 = 0;
  if (lane < rows) {
    const int64_t row = perm ? (int64_t)perm[t0 + lane] : t0 + lane;
    b0 = rowptr[row];
    b1 = rowptr[row + 1];
  }
  VT* base = sval + sptr[s];
  const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  double (*tl)[KT + 1] = tile[w];
  bool bad = false;
  for (int g0 = 0; g0 < rows; g0 += G) {
    int32_t len = 0;  // lane r < G: kept entries of row g0 + r
    for (int j0 = g0; j0 < g0 + G; j0 += 4) {
      int32_t q0[4], q1[4], k[4];
      int32_t longest = 0;
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        q0[u] = __shfl(b0, j0 + u, 64);
        q1[u] = __shfl(b1, j0 + u, 64);
        k[u] = 0;
        longest = max(longest, q1[u] - q0[u]);
      }
      for (int32_t c = 0; c < longest; c += 64) {
        double v[4];
        bool in[4], keep[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int32_t q = q0[u] + c + lane;
          in[u] = q < q1[u];
          v[u] = in[u] ? val64[q] : 0.0;
          keep[u] = in[u] && ((kept[q >> 5] >> (q & 31)) & 1u);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const uint64_t m = __ballot(keep[u]);
          if (keep[u]) {
            const int32_t at = k[u] + __popcll(m & below);
            if (at < KT) tl[j0 - g0 + u][at] = v[u];
            else base[j0 + u + (int64_t)at * kSlice] = (VT)v[u];
          } else if (in[u] && v[u] != 0.0) {
            bad = true;
          }
          k[u] += __popcll(m);
        }
      }
#pragma unroll
      for (int u = 0; u < 4; ++u)
        if (lane == j0 - g0 + u) len = k[u];
    }
    // LDS operations of one wavefront complete in issue order
    __builtin_amdgcn_wave_barrier();
    int32_t top = min(len, KT);
    for (int d = 1; d < G; d <<= 1) top = max(top, __shfl_xor(top, d, 64));
    top = __shfl(top, 0, 64);
    const int r = lane & (G - 1);
    const int32_t mine = min(__shfl(len, r, 64), KT);
    for (int32_t kb = 0; kb < top; kb += kSlice / G) {
      const int32_t at = kb + (lane >> 3);
      if (at < mine) base[g0 + r + (int64_t)at * kSlice] = (VT)tl[r][at];
    }
    __builtin_amdgcn_wave_barrier();
  }
  if (bad) flags[1] = 1;
}

int csr_asm_values(gh_ctx* ctx, gh_csr* A, double** val64) {
  GH_ARG(A && A->ncomp == 1 && !A->host_only,
         "device assembly: needs a scalar matrix with a device layout");
  if (A->prec == GH_F64) {
    *val64 = (double*)A->val;
    return GH_OK;
  }
  if (!A->d_val64) {
    if (hipMalloc((void**)&A->d_val64, sizeof(double) * std::max<int64_t>(A->nnz, 1)) != hipSuccess)
      return fail(GH_ENOMEM, "device assembly: FP64 value buffer");
  }
  *val64 = A->d_val64;
  (void)ctx;
  return GH_OK;
}

int csr_asm_commit(gh_ctx* ctx, gh_csr* A, const char* who) {
  const int64_t n = A->n, nnz = A->nnz;
  double* v64 = nullptr;
  GH_TRY(csr_asm_values(ctx, A, &v64));
  if (!A->d_flags) {
    if (hipMalloc((void**)&A->d_flags, sizeof(int) * 4) != hipSuccess)
      return fail(GH_ENOMEM, "%s: flags", who);
    GH_HIP(hipMemsetAsync(A->d_flags, 0, sizeof(int) * 4, ctx->stream));
  }
  // host staging of the uploads below, alive until the stream is synchronised
  std::vector<int32_t> diag;
  std::vector<uint32_t> kept;
  if (!A->d_diag || !A->d_kept) {
    // from the host copy, which still holds the values of creation here
    if (!A->h_val_ok) return fail(GH_EINVAL, "%s: layout data lost", who);
    diag.assign(n, -1);
    for (int64_t i = 0; i < n; ++i)
      for (int32_t q = A->h_rowptr[i]; q < A->h_rowptr[i + 1]; ++q)
        if (A->h_col[q] == i && diag[i] < 0) diag[i] = q;
    kept.assign((nnz + 31) / 32 + 1, 0u);
    for (int64_t q = 0; q < nnz; ++q)
      if (A->h_val[q] != 0.0) kept[q >> 5] |= 1u << (q & 31);
    GH_TRY(upload(ctx, &A->d_diag, diag.data(), n));
    GH_TRY(upload(ctx, &A->d_kept, kept.data(), (int64_t)kept.size()));
  }
  // the assembly has written v64 on the stream; everything below follows it
  if (A->prec == GH_F64)
    GH_TRY(launch_block_inverse(ctx, n, 1, A->d_diag, v64, (double*)A->dinv, A->d_flags));
  else
    GH_TRY(launch_block_inverse(ctx, n, 1, A->d_diag, v64, (float*)A->dinv, A->d_flags));
  if (A->prec == GH_F32)
    hipLaunchKernelGGL((k_convert<float, double>), dim3(grid_stride(nnz)), dim3(kBlock), 0,
                       ctx->stream, nnz, (float*)A->val, (const double*)v64);
  else if (A->prec == GH_F16)
    hipLaunchKernelGGL((k_convert<__half, double>), dim3(grid_stride(nnz)), dim3(kBlock), 0,
                       ctx->stream, nnz, (__half*)A->val, (const double*)v64);
  GH_HIP(hipGetLastError());
  const Sell& S = A->sell;
  const int64_t nslices = (n + kSlice - 1) / kSlice;
  const dim3 g((unsigned)std::max<int64_t>((nslices + kBlock / kSlice - 1) / (kBlock / kSlice), 1));
  if (A->prec == GH_F64)
    hipLaunchKernelGGL(k_sell_refresh<double>, g, dim3(kBlock), 0, ctx->stream, n, S.sptr, S.perm,
                       A->rowptr, A->d_kept, v64, (double*)S.val, A->d_flags);
  else if (A->prec == GH_F32)
    hipLaunchKernelGGL(k_sell_refresh<float>, g, dim3(kBlock), 0, ctx->stream, n, S.sptr, S.perm,
                       A->rowptr, A->d_kept, v64, (float*)S.val, A->d_flags);
  else
    hipLaunchKernelGGL(k_sell_refresh<__half>, g, dim3(kBlock), 0, ctx->stream, n, S.sptr, S.perm,
                       A->rowptr, A->d_kept, v64, (__half*)S.val, A->d_flags);
  GH_HIP(hipGetLastError());
  int flags[4] = {0, 0, 0, 0};
  GH_HIP(hipMemcpyAsync(flags, A->d_flags, sizeof(flags), hipMemcpyDeviceToHost, ctx->stream));
  GH_HIP(hipStreamSynchronize(ctx->stream));
  A->h_val_ok = false;
  if (flags[0] || flags[1] || flags[2]) {
    GH_HIP(hipMemsetAsync(A->d_flags, 0, sizeof(int) * 4, ctx->stream));
    GH_HIP(hipStreamSynchronize(ctx->stream));
    if (flags[2])
      return fail(GH_EINVAL, "%s: an element coupling has no entry in the matrix pattern", who);
    if (flags[1])
      return fail(GH_EINVAL,
                  "%s: an entry that was zero at creation (left out of the device layout) "
                  "became nonzero",
                  who);
    return fail(GH_ESINGULAR, "%s: singular diagonal", who);
  }
  return GH_OK;
}

int mask_node_flags(gh_ctx* ctx, gh_mask* M, const uint8_t** out) {
  if (!M->d_nodeflag) {
    std::vector<uint8_t> f(std::max<int64_t>(M->n, 1), 0);
    for (int32_t v : M->h_nodes) f[v] = 1;
    GH_TRY(upload(ctx, &M->d_nodeflag, f.data(), (int64_t)f.size()));
    GH_HIP(hipStreamSynchronize(ctx->stream));  // f is host staging
  }
  *out = M->d_nodeflag;
  return GH_OK;
}

int csr_create_impl(gh_ctx* ctx, int64_t n, int ncomp, const int32_t* rowptr, const int32_t* col,
                    const double* val, int prec, bool check_sorted, gh_csr** out) {
  GH_ARG(ctx && out && rowptr && n >= 0 && ncomp >= 1 && ncomp <= 8,
         "gh_csr_create: bad arguments (ncomp must be 1..8)");
  GH_ARG(prec == GH_F64 || prec == GH_F32 || prec == GH_F16, "gh_csr_create: bad precision");
  GH_ARG(n < (int64_t)kIdxMask, "gh_csr_create: too many rows");
  *out = nullptr;
  const int64_t nnz = rowptr[n];
  GH_ARG(rowptr[0] == 0 && nnz >= 0 && (nnz == 0 || (col && val)), "gh_csr_create: bad rowptr");
  const int BB = ncomp * ncomp;
  std::vector<int32_t> diag(n);
  for (int64_t i = 0; i < n; ++i)
    if (rowptr[i + 1] < rowptr[i]) return fail(GH_EINVAL, "gh_csr_create: rowptr not monotone");
  // per row (on the host threads): 1 column out of range, 2 columns not
  // ascending, 3 no diagonal; the first faulty row is reported
  std::vector<uint8_t> bad(n, 0);
  parallel_rows(n, [&](int64_t a, int64_t b) {
    for (int64_t i = a; i < b; ++i) {
      diag[i] = -1;
      for (int32_t p = rowptr[i]; p < rowptr[i + 1] && !bad[i]; ++p) {
        if (col[p] < 0 || col[p] >= n) bad[i] = 1;
        else if (check_sorted && p > rowptr[i] && col[p] <= col[p - 1]) bad[i] = 2;
        else if (col[p] == i && diag[i] < 0) diag[i] = p;
      }
      if (!bad[i] && diag[i] < 0) bad[i] = 3;
    }
  });
  for (int64_t i = 0; i < n; ++i) {
    if (bad[i] == 1)
      for (int32_t p = rowptr[i]; p < rowptr[i + 1]; ++p)
        if (col[p] < 0 || col[p] >= n)
          return fail(GH_EINVAL, "gh_csr_create: column %d out of range in row %lld", col[p],
                      (long long)i);
    if (bad[i] == 2)
      return fail(GH_EINVAL, "gh_csr_create: columns of row %lld not strictly ascending",
                  (long long)i);
    if (bad[i] == 3) return fail(GH_EINVAL, "gh_csr_create: row %lld has no diagonal", (long long)i);
  }
  gh_csr* A = new gh_csr();
  A->ctx = ctx;
  A->n = n;
  A->ncomp = ncomp;
  A->nnz = nnz;
  A->prec = prec;
  A->h_rowptr.assign(rowptr, rowptr + n + 1);
  A->h_col.assign(col, col + nnz);
  A->h_val.assign(val, val + nnz * BB);
  int rc = upload(ctx, &A->rowptr, rowptr, n + 1);
  if (!rc) rc = upload(ctx, &A->col, col, nnz);
  double* dval64 = nullptr;
  if (!rc) rc = upload(ctx, &dval64, val, nnz * BB);
  int32_t* ddiag = nullptr;
  if (!rc) rc = upload(ctx, &ddiag, diag.data(), n);
  int* dsing = nullptr;
  if (!rc && hipMalloc((void**)&dsing, sizeof(int)) != hipSuccess) rc = fail(GH_ENOMEM, "alloc");
  if (!rc && hipMemsetAsync(dsing, 0, sizeof(int), ctx->stream) != hipSuccess) rc = GH_EHIP;
  // inverse diagonal blocks, FP64 then narrowed for the FP32 paths
  const size_t dsz = (size_t)std::max<int64_t>(n * BB, 1);
  if (!rc) {
    if (hipMalloc(&A->dinv, dsz * (prec == GH_F64 ? 8 : 4)) != hipSuccess)
      rc = fail(GH_ENOMEM, "alloc dinv");
    else if (prec == GH_F64)
      rc = launch_block_inverse(ctx, n, ncomp, ddiag, dval64, (double*)A->dinv, dsing);
    else
      rc = launch_block_inverse(ctx, n, ncomp, ddiag, dval64, (float*)A->dinv, dsing);
  }
  if (!rc) {
    if (prec == GH_F64) {
      A->val = dval64;
      dval64 = nullptr;
    } else {
      const size_t m = (size_t)std::max<int64_t>(nnz * BB, 1);
      if (hipMalloc(&A->val, m * (prec == GH_F32 ? 4 : 2)) != hipSuccess)
        rc = fail(GH_ENOMEM, "alloc val");
      else if (prec == GH_F32)
        hipLaunchKernelGGL((k_convert<float, double>), dim3(grid_stride(nnz * BB)), dim3(kBlock),
                           0, ctx->stream, nnz * BB, (float*)A->val, (const double*)dval64);
      else
        hipLaunchKernelGGL((k_convert<__half, double>), dim3(grid_stride(nnz * BB)),
                           dim3(kBlock), 0, ctx->stream, nnz * BB, (__half*)A->val,
                           (const double*)dval64);
    }
  }
  if (!rc && ncomp == 1) {
    rc = sell_build(ctx, n, rowptr, reinterpret_cast<const uint32_t*>(col), val, prec, &A->sell);
  }
  if (!rc && ncomp >= 2 && ncomp <= 4) rc = bsell_build(ctx, n, ncomp, rowptr, col, val, prec, &A->bsell);
  int sing = 0;
  if (!rc && hipMemcpyAsync(&sing, dsing, sizeof(int), hipMemcpyDeviceToHost, ctx->stream) != hipSuccess)
    rc = GH_EHIP;
  if (!rc && hipStreamSynchronize(ctx->stream) != hipSuccess) rc = fail(GH_EHIP, "sync");
  if (dval64) (void)hipFree(dval64);
  if (ddiag) (void)hipFree(ddiag);
  if (dsing) (void)hipFree(dsing);
  if (!rc && sing) rc = fail(GH_ESINGULAR, "gh_csr_create: singular diagonal block");
  if (rc) {
    ctx_retain(ctx);  // balanced by gh_csr_destroy
    gh_csr_destroy(A);
    return rc;
  }
  ctx_retain(ctx);
  *out = A;
  return GH_OK;
}

}  // namespace gh

extern "C" {

int gh_csr_create(gh_ctx* ctx, int64_t n, int ncomp, const int32_t* rowptr, const int32_t* col,
                  const double* val, int prec, gh_csr** out) {
  return csr_create_impl(ctx, n, ncomp, rowptr, col, val, prec, true, out);
}

int gh_csr_create_host(gh_ctx* ctx, int64_t n, int ncomp, const int32_t* rowptr,
                       const int32_t* col, const double* val, gh_csr** out) {
  GH_ARG(ctx && out && rowptr && n >= 0 && ncomp >= 1 && ncomp <= 8,
         "gh_csr_create_host: bad arguments (ncomp must be 1..8)");
  GH_ARG(n < (int64_t)kIdxMask, "gh_csr_create_host: too many rows");
  *out = nullptr;
  const int64_t nnz = rowptr[n];
  GH_ARG(rowptr[0] == 0 && nnz >= 0 && (nnz == 0 || (col && val)),
         "gh_csr_create_host: bad rowptr");
  for (int64_t i = 0; i < n; ++i) {
    if (rowptr[i + 1] < rowptr[i]) return fail(GH_EINVAL, "gh_csr_create_host: rowptr not monotone");
    bool diag = false;
    for (int32_t p = rowptr[i]; p < rowptr[i + 1]; ++p) {
      if (col[p] < 0 || col[p] >= n)
        return fail(GH_EINVAL, "gh_csr_create_host: column %d out of range in row %lld", col[p],
                    (long long)i);
      if (p > rowptr[i] && col[p] <= col[p - 1])
        return fail(GH_EINVAL, "gh_csr_create_host: columns of row %lld not strictly ascending",
                    (long long)i);
      diag = diag || col[p] == i;
    }
    if (!diag) return fail(GH_EINVAL, "gh_csr_create_host: row %lld has no diagonal", (long long)i);
  }
  gh_csr* A = new gh_csr();
  A->ctx = ctx;
  A->n = n;
  A->ncomp = ncomp;
  A->nnz = nnz;
  A->prec = GH_F64;
  A->host_only = true;
  A->h_rowptr.assign(rowptr, rowptr + n + 1);
  A->h_col.assign(col, col + nnz);
  A->h_val.assign(val, val + nnz * (int64_t)ncomp * ncomp);
  ctx_retain(ctx);  // balanced by gh_csr_destroy
  *out = A;
  return GH_OK;
}

int gh_csr_destroy(gh_csr* A) {
  if (!A) return GH_OK;
  gh_ctx* ctx = A->ctx;
  if (ctx && ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  for (void* p : {(void*)A->rowptr, (void*)A->col, A->val, A->dinv, (void*)A->t_rowptr,
                  (void*)A->t_col, (void*)A->t_pos, (void*)A->d_diag, (void*)A->d_kept,
                  (void*)A->d_val64, (void*)A->d_flags})
    if (p) (void)hipFree(p);
  sell_free(&A->sell);
  bsell_free(&A->bsell);
  delete A;
  ctx_release(ctx);
  return GH_OK;
}

int64_t gh_csr_n(const gh_csr* A) { return A ? A->n : -1; }
int64_t gh_csr_nnz(const gh_csr* A) { return A ? A->nnz : -1; }
int gh_csr_prec(const gh_csr* A) { return A ? A->prec : -1; }
int gh_csr_layout(const gh_csr* A, int64_t* out) {
  GH_ARG(A && out, "gh_csr_layout: NULL argument");
  GH_ARG(!A->host_only, "gh_csr_layout: the matrix has no device layout (created by gh_csr_create_host)");
  if (A->ncomp == 1) {
    out[0] = A->sell.stored;
    out[1] = A->sell.slots;
    out[2] = A->sell.nslices;
    out[3] = A->sell.narrow;
    out[4] = A->sell.patterned;
    out[5] = A->sell.ntab;
    out[6] = A->sell.col_bytes;
    out[7] = A->sell.perm != nullptr;
  } else if (A->bsell.B) {
    // component rows: entries are block rows (B values each), 32-bit columns
    out[0] = A->bsell.stored;
    out[1] = A->bsell.slots;
    out[2] = A->bsell.nslices;
    out[3] = out[4] = out[5] = 0;
    out[6] = A->bsell.stored * 4;
    out[7] = 1;
  } else {
    out[0] = out[1] = A->nnz;
    out[2] = out[3] = out[4] = out[5] = 0;
    out[6] = A->nnz * 4;
    out[7] = 0;
  }
  return GH_OK;
}

static int vec_matches(const gh_csr* A, const gh_vec* v) {
  return v && v->n == A->n && v->ncomp == A->ncomp;
}

int gh_spmv(gh_ctx* ctx, const gh_csr* A, gh_vec* y, const gh_vec* x, double d, double b) {
  GH_ARG(ctx && A && vec_matches(A, y) && vec_matches(A, x), "gh_spmv: shape mismatch");
  GH_ARG(!A->host_only, "gh_spmv: the matrix has no device layout (created by gh_csr_create_host)");
  GH_ARG(y->d != x->d, "gh_spmv: x and y must not alias");
  return csr_spmv(ctx, A, y->prec, y->d, x->prec, x->d, d, b);
}

int gh_residual(gh_ctx* ctx, const gh_csr* A, gh_vec* r, const gh_vec* x, const gh_vec* b) {
  GH_ARG(ctx && A && vec_matches(A, r) && vec_matches(A, x) && vec_matches(A, b),
         "gh_residual: shape mismatch");
  GH_ARG(!A->host_only, "gh_residual: the matrix has no device layout (created by gh_csr_create_host)");
  GH_ARG(r->prec == x->prec && x->prec == b->prec, "gh_residual: precision mismatch");
  GH_ARG(r->d != x->d, "gh_residual: r and x must not alias");
  if (r->prec == GH_F64 && A->prec != GH_F64)
    return fail(GH_EINVAL, "gh_residual: FP64 vectors need an FP64 matrix");
  return csr_residual(ctx, A, r->prec, r->d, x->d, b->d);
}

int gh_smooth(gh_ctx* ctx, const gh_csr* A, gh_vec* x, const gh_vec* b, gh_vec* h, int niter,
              double omega, int smoother) {
  GH_ARG(ctx && A && vec_matches(A, x) && vec_matches(A, b) && vec_matches(A, h),
         "gh_smooth: shape mismatch");
  GH_ARG(!A->host_only, "gh_smooth: the matrix has no device layout (created by gh_csr_create_host)");
  GH_ARG(x->prec == b->prec && b->prec == h->prec, "gh_smooth: precision mismatch");
  GH_ARG(smoother == 0 || smoother == 1, "gh_smooth: smoother must be 0 (jacobi) or 1 (richardson)");
  void* cur = x->d;
  void* nxt = h->d;
  for (int it = 0; it < niter; ++it) {
    GH_TRY(csr_jacobi_sweep(ctx, A, x->prec, nxt, cur, b->d, omega, smoother, false));
    std::swap(cur, nxt);
  }
  if (cur != x->d) return vec_convert(ctx, x->size, x->prec, x->d, x->prec, cur);
  return GH_OK;
}

int gh_jacobi(gh_ctx* ctx, const gh_csr* A, gh_vec* h) {
  GH_ARG(ctx && A && vec_matches(A, h), "gh_jacobi: shape mismatch");
  GH_ARG(!A->host_only, "gh_jacobi: the matrix has no device layout (created by gh_csr_create_host)");
  // h = D^{-1} h, as a sweep from x = 0 with omega = 1 into a temporary
  gh_vec t;
  GH_TRY(vec_alloc(ctx, h->n, h->ncomp, h->prec, &t));
  int rc = csr_jacobi_sweep(ctx, A, h->prec, t.d, nullptr, h->d, 1.0, 0, true);
  if (!rc) rc = vec_convert(ctx, h->size, h->prec, h->d, h->prec, t.d);
  (void)hipStreamSynchronize(ctx->stream);
  vec_free(&t);
  return rc;
}

int gh_jacobi_noalloc(gh_ctx* ctx, const gh_csr* A, gh_vec* h) {
  GH_ARG(ctx && A && vec_matches(A, h), "gh_jacobi_noalloc: shape mismatch");
  GH_ARG(!A->host_only,
         "gh_jacobi_noalloc: the matrix has no device layout (created by gh_csr_create_host)");
  const size_t bytes = (size_t)h->size * prec_bytes(h->prec);
  if (bytes > kNoAllocScratchBytes)
    return fail(GH_EINVAL,
                "gh_jacobi_noalloc: h holds %zu bytes, more than its %zu-byte scratch (use gh_jacobi)",
                bytes, kNoAllocScratchBytes);
  if (!ctx->noalloc_scratch)
    return fail(GH_EHIP, "gh_jacobi_noalloc: no scratch address for device %d", ctx->device);
  // the same two launches as gh_jacobi, with the scratch as the temporary: the
  // sweep from zero writes every entry of it, so nothing needs clearing
  GH_TRY(csr_jacobi_sweep(ctx, A, h->prec, ctx->noalloc_scratch, nullptr, h->d, 1.0, 0, true));
  return vec_convert(ctx, h->size, h->prec, h->d, h->prec, ctx->noalloc_scratch);
}

// ---------------------------------------------------------------- transfer
int gh_transfer_create(gh_ctx* ctx, int64_t ncoarse, const int32_t* c2f, int64_t n2,
                       const int32_t* zweier, int64_t n4, const int32_t* vierer, int64_t n8,
                       const int32_t* achter, gh_transfer** out) {
  GH_ARG(ctx && out && ncoarse >= 0 && c2f, "gh_transfer_create: bad arguments");
  *out = nullptr;
  int64_t nfine = -1;
  for (int64_t i = 0; i < ncoarse; ++i) nfine = std::max<int64_t>(nfine, c2f[i]);
  for (int64_t i = 0; i < n2; ++i) nfine = std::max<int64_t>(nfine, zweier[3 * i]);
  for (int64_t i = 0; i < n4; ++i) nfine = std::max<int64_t>(nfine, vierer[5 * i]);
  for (int64_t i = 0; i < n8; ++i) nfine = std::max<int64_t>(nfine, achter[9 * i]);
  nfine++;
  GH_ARG(nfine <= (int64_t)kIdxMask, "gh_transfer_create: too many nodes for packed columns");
  // weight code of each fine node: I2 -> 1, I4 -> 2, I8 -> 3, else 0
  // (lookup order of MgInterpolatorMatrix::BasicInit, mginterpolatormatrix.cc:97-121)
  std::vector<uint8_t> code(nfine, 0);
  for (int64_t i = n8 - 1; i >= 0; --i) code[achter[9 * i]] = 3;
  for (int64_t i = n4 - 1; i >= 0; --i) code[vierer[5 * i]] = 2;
  for (int64_t i = n2 - 1; i >= 0; --i) code[zweier[3 * i]] = 1;
  std::vector<std::pair<int32_t, int32_t>> F, C;  // (row, col)
  const int64_t np = ncoarse + 2 * n2 + 4 * n4 + 8 * n8;
  F.reserve(np);
  C.reserve(np);
  for (int64_t i = 0; i < ncoarse; ++i) {
    GH_ARG(c2f[i] >= 0, "gh_transfer_create: negative c2f entry");
    F.emplace_back(c2f[i], (int32_t)i);
    C.emplace_back((int32_t)i, c2f[i]);
  }
  auto add = [&](const int32_t* a, int64_t cnt, int w) -> bool {
    for (int64_t i = 0; i < cnt; ++i)
      for (int j = 0; j < w; ++j) {
        int32_t f = a[(w + 1) * i], c = a[(w + 1) * i + 1 + j];
        if (c < 0 || c >= ncoarse) return false;
        F.emplace_back(f, c);
        C.emplace_back(c, f);
      }
    return true;
  };
  GH_ARG(add(zweier, n2, 2) && add(vierer, n4, 4) && add(achter, n8, 8),
         "gh_transfer_create: parent index out of range");
  std::sort(F.begin(), F.end());
  F.erase(std::unique(F.begin(), F.end()), F.end());
  std::sort(C.begin(), C.end());
  C.erase(std::unique(C.begin(), C.end()), C.end());
  gh_transfer* T = new gh_transfer();
  T->ctx = ctx;
  T->ncoarse = ncoarse;
  T->nfine = nfine;
  T->h_rrow.assign(ncoarse + 1, 0);
  T->h_prow.assign(nfine + 1, 0);
  for (auto& e : C) {
    T->h_rrow[e.first + 1]++;
    T->h_rcol.push_back((uint32_t)e.second | ((uint32_t)code[e.second] << 29));
  }
  for (auto& e : F) {
    T->h_prow[e.first + 1]++;
    T->h_pcol.push_back((uint32_t)e.second | ((uint32_t)code[e.first] << 29));
  }
  for (int64_t i = 0; i < ncoarse; ++i) T->h_rrow[i + 1] += T->h_rrow[i];
  for (int64_t i = 0; i < nfine; ++i) T->h_prow[i + 1] += T->h_prow[i];
  int rc = upload(ctx, &T->r_row, T->h_rrow.data(), T->h_rrow.size());
  if (!rc) rc = upload(ctx, &T->r_col, T->h_rcol.data(), T->h_rcol.size());
  if (!rc) rc = upload(ctx, &T->p_row, T->h_prow.data(), T->h_prow.size());
  if (!rc) rc = upload(ctx, &T->p_col, T->h_pcol.data(), T->h_pcol.size());
  if (!rc) rc = sell_build(ctx, ncoarse, T->h_rrow.data(), T->h_rcol.data(), nullptr, GH_F64, &T->r_sell);
  if (!rc) rc = sell_build(ctx, nfine, T->h_prow.data(), T->h_pcol.data(), nullptr, GH_F64, &T->p_sell);
  if (!rc && hipStreamSynchronize(ctx->stream) != hipSuccess) rc = fail(GH_EHIP, "sync");
  if (rc) {
    ctx_retain(ctx);  // balanced by gh_transfer_destroy
    gh_transfer_destroy(T);
    return rc;
  }
  ctx_retain(ctx);
  *out = T;
  return GH_OK;
}

int gh_transfer_destroy(gh_transfer* T) {
  if (!T) return GH_OK;
  gh_ctx* ctx = T->ctx;
  if (ctx && ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  for (void* p : {(void*)T->r_row, (void*)T->r_col, (void*)T->p_row, (void*)T->p_col})
    if (p) (void)hipFree(p);
  sell_free(&T->r_sell);
  sell_free(&T->p_sell);
  delete T;
  ctx_release(ctx);
  return GH_OK;
}

int64_t gh_transfer_nfine(const gh_transfer* T) { return T ? T->nfine : -1; }

int gh_restrict_zero(gh_ctx* ctx, const gh_transfer* T, gh_vec* uc, const gh_vec* uf) {
  GH_ARG(ctx && T && uc && uf && uc->n == T->ncoarse && uf->n == T->nfine &&
             uc->ncomp == uf->ncomp && uc->prec == uf->prec,
         "gh_restrict_zero: shape mismatch");
  return transfer_restrict(ctx, T, uc->ncomp, uc->prec, uc->d, uf->d, nullptr);
}

int gh_prolongate_add(gh_ctx* ctx, const gh_transfer* T, gh_vec* uf, const gh_vec* uc) {
  GH_ARG(ctx && T && uc && uf && uc->n == T->ncoarse && uf->n == T->nfine &&
             uc->ncomp == uf->ncomp && uc->prec == uf->prec,
         "gh_prolongate_add: shape mismatch");
  return transfer_prolongate_add(ctx, T, uf->ncomp, uf->prec, uf->d, uc->d);
}

// ------------------------------------------------------------ hanging nodes
int gh_hn_create(gh_ctx* ctx, int kind, int64_t n, int64_t ne, const int32_t* edges,
                 int64_t nf, const int32_t* faces, gh_hn** out) {
  GH_ARG(ctx && out && (kind == 12 || kind == 13 || kind == 22 || kind == 23),
         "gh_hn_create: kind must be 12, 13, 22 or 23");
  GH_ARG(ne >= 0 && nf >= 0 && (ne == 0 || edges) && (nf == 0 || faces),
         "gh_hn_create: bad lists");
  GH_ARG(!((kind == 12 || kind == 22) && nf > 0), "gh_hn_create: 2D structures have no faces");
  *out = nullptr;
  const bool q2 = (kind == 22 || kind == 23);
  // hanging set and parent check: the gather-form kernels need regular parents
  std::vector<uint8_t> hang(n, 0);
  auto chk = [&](int32_t v) { return v >= 0 && v < n; };
  for (int64_t k = 0; k < ne; ++k) {
    GH_ARG(chk(edges[4 * k]), "gh_hn_create: node out of range");
    GH_ARG(!hang[edges[4 * k]], "gh_hn_create: hanging node listed twice");
    hang[edges[4 * k]] = 1;
  }
  for (int64_t k = 0; k < nf; ++k) {
    GH_ARG(chk(faces[10 * k]), "gh_hn_create: node out of range");
    GH_ARG(!hang[faces[10 * k]], "gh_hn_create: hanging node listed twice");
    hang[faces[10 * k]] = 1;
  }
  // contributions (parent, src, w) in the reference's sequential order
  struct Contr {
    int32_t p, h;
    double w;
  };
  std::vector<Contr> cs;
  const double q1w[2] = {0.5, 0.5}, q2w[3] = {0.375, 0.75, -0.125};
  auto edge_contr = [&]() -> bool {
    for (int64_t k = 0; k < ne; ++k) {
      const int32_t* e = edges + 4 * k;
      for (int j = 0; j < (q2 ? 3 : 2); ++j) {
        if (!chk(e[1 + j]) || hang[e[1 + j]]) return false;
        cs.push_back({e[1 + j], e[0], q2 ? q2w[j] : q1w[j]});
      }
    }
    return true;
  };
  auto face_contr = [&]() -> bool {
    for (int64_t k = 0; k < nf; ++k) {
      const int32_t* f = faces + 10 * k;
      if (kind == 23) {
        for (int j = 0; j < 9; ++j) {
          if (!chk(f[1 + j]) || hang[f[1 + j]]) return false;
          cs.push_back({f[1 + j], f[0], q2w[j / 3] * q2w[j % 3]});
        }
      } else {
        const int sel[4] = {0, 1, 3, 4};
        for (int j : sel) {
          if (!chk(f[1 + j]) || hang[f[1 + j]]) return false;
          cs.push_back({f[1 + j], f[0], 0.25});
        }
      }
    }
    return true;
  };
  bool ok;
  if (kind == 23)
    ok = face_contr() && edge_contr();  // HNStructureQ23d: faces, then edges
  else
    ok = edge_contr() && face_contr();  // HNStructureQ13d: edges, then faces
  GH_ARG(ok, "gh_hn_create: a parent is out of range or itself hanging");
  // stable grouping by parent keeps the per-parent order
  std::vector<int64_t> order(cs.size());
  for (size_t i = 0; i < cs.size(); ++i) order[i] = (int64_t)i;
  std::stable_sort(order.begin(), order.end(),
                   [&](int64_t a, int64_t b) { return cs[a].p < cs[b].p; });
  std::vector<int32_t> parent, row(1, 0), src;
  std::vector<double> w;
  for (size_t i = 0; i < order.size(); ++i) {
    const Contr& c = cs[order[i]];
    if (parent.empty() || parent.back() != c.p) {
      if (!parent.empty()) row.push_back((int32_t)src.size());
      parent.push_back(c.p);
    }
    src.push_back(c.h);
    w.push_back(c.w);
  }
  if (!parent.empty()) row.push_back((int32_t)src.size());
  std::vector<int32_t> hanging;
  for (int64_t k = 0; k < ne; ++k) hanging.push_back(edges[4 * k]);
  for (int64_t k = 0; k < nf; ++k) hanging.push_back(faces[10 * k]);

  gh_hn* H = new gh_hn();
  H->ctx = ctx;
  H->kind = kind;
  H->n = n;
  H->ne = ne;
  H->nf = nf;
  H->np = (int64_t)parent.size();
  H->h_edges.assign(edges, edges + 4 * ne);
  H->h_faces.assign(faces, faces + 10 * nf);
  int rc = upload(ctx, &H->edges, edges, 4 * ne);
  if (!rc) rc = upload(ctx, &H->faces, faces, 10 * nf);
  if (!rc) rc = upload(ctx, &H->d_parent, parent.data(), parent.size());
  if (!rc) rc = upload(ctx, &H->d_row, row.data(), row.size());
  if (!rc) rc = upload(ctx, &H->d_src, src.data(), src.size());
  if (!rc) rc = upload(ctx, &H->d_w, w.data(), w.size());
  if (!rc) rc = upload(ctx, &H->hanging, hanging.data(), hanging.size());
  if (!rc && hipStreamSynchronize(ctx->stream) != hipSuccess) rc = fail(GH_EHIP, "sync");
  if (rc) {
    ctx_retain(ctx);  // balanced by gh_hn_destroy
    gh_hn_destroy(H);
    return rc;
  }
  ctx_retain(ctx);
  *out = H;
  return GH_OK;
}

int gh_hn_destroy(gh_hn* H) {
  if (!H) return GH_OK;
  gh_ctx* ctx = H->ctx;
  if (ctx && ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  for (void* p : {(void*)H->edges, (void*)H->faces, (void*)H->d_parent, (void*)H->d_row,
                  (void*)H->d_src, (void*)H->d_w, (void*)H->hanging})
    if (p) (void)hipFree(p);
  delete H;
  ctx_release(ctx);
  return GH_OK;
}

static int hn_check(const gh_hn* H, const gh_vec* x) { return H && x && x->n == H->n; }

int gh_hn_average(gh_ctx* ctx, const gh_hn* H, gh_vec* x) {
  GH_ARG(ctx && hn_check(H, x), "gh_hn_average: shape mismatch");
  return hn_apply(ctx, H, x->ncomp, x->prec, x->d, 0);
}
int gh_hn_distribute(gh_ctx* ctx, const gh_hn* H, gh_vec* x) {
  GH_ARG(ctx && hn_check(H, x), "gh_hn_distribute: shape mismatch");
  return hn_apply(ctx, H, x->ncomp, x->prec, x->d, 1);
}
int gh_hn_zero(gh_ctx* ctx, const gh_hn* H, gh_vec* x) {
  GH_ARG(ctx && hn_check(H, x), "gh_hn_zero: shape mismatch");
  return hn_apply(ctx, H, x->ncomp, x->prec, x->d, 2);
}

// ---------------------------------------------------------------- Dirichlet
int gh_mask_create(gh_ctx* ctx, int64_t n, int ncomp, int64_t nidx, const int32_t* idx,
                   int ncomps, const int32_t* comps, gh_mask** out) {
  GH_ARG(ctx && out && n >= 0 && ncomp >= 1 && nidx >= 0 && ncomps >= 0 &&
             (nidx == 0 || idx) && (ncomps == 0 || comps),
         "gh_mask_create: bad arguments");
  *out = nullptr;
  std::vector<int32_t> flat;
  flat.reserve(nidx * ncomps);
  for (int64_t i = 0; i < nidx; ++i) {
    GH_ARG(idx[i] >= 0 && idx[i] < n, "gh_mask_create: node out of range");
    for (int k = 0; k < ncomps; ++k) {
      GH_ARG(comps[k] >= 0 && comps[k] < ncomp, "gh_mask_create: component out of range");
      flat.push_back((int32_t)(idx[i] * ncomp + comps[k]));
    }
  }
  gh_mask* M = new gh_mask();
  M->ctx = ctx;
  M->n = n;
  M->ncomp = ncomp;
  M->nidx = nidx;
  M->ncomps = ncomps;
  M->h_nodes.assign(idx, idx + nidx);
  M->h_comps.assign(comps, comps + ncomps);
  int rc = upload(ctx, &M->idx, flat.data(), flat.size());
  if (!rc && hipStreamSynchronize(ctx->stream) != hipSuccess) rc = fail(GH_EHIP, "sync");
  if (rc) {
    ctx_retain(ctx);  // balanced by gh_mask_destroy
    gh_mask_destroy(M);
    return rc;
  }
  ctx_retain(ctx);
  *out = M;
  return GH_OK;
}

int gh_mask_destroy(gh_mask* M) {
  if (!M) return GH_OK;
  gh_ctx* ctx = M->ctx;
  if (ctx && ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  if (M->idx) (void)hipFree(M->idx);
  if (M->d_nodeflag) (void)hipFree(M->d_nodeflag);
  delete M;
  ctx_release(ctx);
  return GH_OK;
}

int gh_mask_zero(gh_ctx* ctx, const gh_mask* M, gh_vec* x) {
  GH_ARG(ctx && M && x && x->n == M->n && x->ncomp == M->ncomp, "gh_mask_zero: shape mismatch");
  return mask_apply(ctx, M, x->prec, x->d);
}

}  // extern "C"
