// ILU(0) smoother of the gascoignelib_amd C ABI: Gascoigne's SparseBlockIlu
// (src/LinAlg/sparseblockilu.xx), the CPU path's default linear_smooth "ilu"
// (src/Solver/solverdata.cc:55), on the device.
//
// Structure (ConstructStructure, :70-127): ILU row i is node perm[i]; its
// columns are the matrix row's columns renumbered by the inverse order and
// sorted ascending.  copy_entries (:131-166) places the matrix blocks there.
// Both are host work at creation.  Factorisation (compute_ilu, :240-275) and
// the triangular solves (forward / backward, :193-222) run on the device by
// level scheduling: row i of the forward DAG waits only for the rows its lower
// part references, so the rows of one level (1 + the largest level they
// reference) are independent and run as one launch; the backward solve uses
// the levels of the upper part (level 0: rows without upper entries).  Each row's operations run
// in one thread in the reference's order, every product and difference
// rounded separately, so the factors and the FP64 solves equal the CPU ILU's
// bit for bit.  FP32 solves (the reduced-precision V-cycle) use an FP32 copy
// of the factors.
#include <algorithm>
#include <functional>
#include <numeric>
#include <vector>

#include "gh_kernels.h"

using namespace gh;

struct gh_ilu {
  gh_ctx* ctx = nullptr;
  int64_t n = 0, nnz = 0;
  int B = 1;
  int32_t* p = nullptr;       // ILU row -> node
  int32_t* q = nullptr;       // node -> ILU row
  int32_t* rowptr = nullptr;
  int32_t* col = nullptr;
  int32_t* diag = nullptr;
  double* val = nullptr;      // factors, nnz blocks of B*B (row-major), FP64
  float* val32 = nullptr;     // the same in FP32
  void* y = nullptr;          // work vector, n*B doubles (or floats)
  int32_t* rows_f = nullptr;  // rows grouped by forward level
  int32_t* rows_b = nullptr;  // rows grouped by backward level
  std::vector<int32_t> lptr_f, lptr_b;  // level boundaries (host)
  int* d_flag = nullptr;
};

namespace {

// compute_ilu for the rows of one forward level
template <int B>
__global__ __launch_bounds__(kBlock) void k_ilu_factor(int32_t cnt, const int32_t* __restrict__ rows,
                                                       const int32_t* __restrict__ rowptr,
                                                       const int32_t* __restrict__ col,
                                                       const int32_t* __restrict__ diag,
                                                       double* __restrict__ val, int* flag) {
  constexpr int BB = B * B;
  const int32_t t = (int32_t)(blockIdx.x * kBlock + threadIdx.x);
  if (t >= cnt) return;
  const int32_t i = rows[t];
  const int32_t di = diag[i], e = rowptr[i + 1];
  for (int32_t pk = rowptr[i]; pk < di; ++pk) {
    const int32_t k = col[pk];
    double* M = val + (int64_t)pk * BB;
    const double* D = val + (int64_t)diag[k] * BB;
    // a(i,k) *= a(k,k)^-1 (FMatrixBlock::operator*=, fmatrixblock.h:174-190)
    double m[BB];
#pragma unroll
    for (int r = 0; r < B; ++r) {
      double vh[B];
#pragma unroll
      for (int c = 0; c < B; ++c) vh[c] = M[r * B + c];
#pragma unroll
      for (int c = 0; c < B; ++c) {
        double a = 0.0;
#pragma unroll
        for (int u = 0; u < B; ++u) a = add_rn(a, mul_rn(vh[u], D[u * B + c]));
        m[r * B + c] = a;
      }
    }
#pragma unroll
    for (int x = 0; x < BB; ++x) M[x] = m[x];
    // a(i,j) -= a(i,k) a(k,j) for the upper entries j of row k (submult)
    for (int32_t pj = diag[k] + 1; pj < rowptr[k + 1]; ++pj) {
      const int32_t j = col[pj];
      for (int32_t ph = rowptr[i]; ph < e; ++ph) {
        if (col[ph] != j) continue;
        double* Hm = val + (int64_t)ph * BB;
        const double* Cm = val + (int64_t)pj * BB;
#pragma unroll
        for (int r = 0; r < B; ++r)
#pragma unroll
          for (int c = 0; c < B; ++c) {
            double h = Hm[r * B + c];
#pragma unroll
            for (int u = 0; u < B; ++u) h = add_rn(h, -mul_rn(m[r * B + u], Cm[u * B + c]));
            Hm[r * B + c] = h;
          }
        break;
      }
    }
  }
  // a(i,i) <- a(i,i)^-1 (FMatrixBlock::inverse = NodeMatrix::gauss_jordan)
  double d[BB];
  double* Dm = val + (int64_t)di * BB;
#pragma unroll
  for (int x = 0; x < BB; ++x) d[x] = Dm[x];
  if (!gauss_jordan<B>(d)) {
    flag[0] = 1;
    return;
  }
#pragma unroll
  for (int x = 0; x < BB; ++x) Dm[x] = d[x];
}

// y_i = x_{order[i]} (hin: yp(i) = x(p[i]); her: x(i) = yp(q[i]))
template <typename T>
__global__ __launch_bounds__(kBlock) void k_ilu_gather(int64_t n, int B, const int32_t* __restrict__ order,
                                                       const T* __restrict__ x, T* __restrict__ y) {
  const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (t >= n * B) return;
  const int64_t i = t / B;
  const int c = (int)(t - i * B);
  y[t] = x[(int64_t)order[i] * B + c];
}

template <int B, typename T, typename VT>
__device__ __forceinline__ void block_subtract(T* yi, const VT* M, const T* yc) {
  // FMatrixBlock::subtract (fmatrixblock.h:136-146): y_i -= M y_c, row by row
#pragma unroll
  for (int r = 0; r < B; ++r) {
    T a = yi[r];
#pragma unroll
    for (int c = 0; c < B; ++c) a = add_rn(a, -mul_rn((T)M[r * B + c], yc[c]));
    yi[r] = a;
  }
}

// forward (:193-203) for the rows of one level
template <int B, typename T, typename VT>
__global__ __launch_bounds__(kBlock) void k_ilu_forward(int32_t cnt, const int32_t* __restrict__ rows,
                                                        const int32_t* __restrict__ rowptr,
                                                        const int32_t* __restrict__ col,
                                                        const int32_t* __restrict__ diag,
                                                        const VT* __restrict__ val, T* __restrict__ y) {
  const int32_t t = (int32_t)(blockIdx.x * kBlock + threadIdx.x);
  if (t >= cnt) return;
  const int32_t i = rows[t];
  T yi[B];
#pragma unroll
  for (int r = 0; r < B; ++r) yi[r] = y[(int64_t)i * B + r];
  for (int32_t pos = rowptr[i]; pos < diag[i]; ++pos) {
    T yc[B];
#pragma unroll
    for (int c = 0; c < B; ++c) yc[c] = y[(int64_t)col[pos] * B + c];
    block_subtract<B>(yi, val + (int64_t)pos * B * B, yc);
  }
#pragma unroll
  for (int r = 0; r < B; ++r) y[(int64_t)i * B + r] = yi[r];
}

// backward (:207-222) for the rows of one level: the upper entries from the
// last to the first, then y_i = a(i,i)^-1 y_i (FMatrixBlock::vmult)
template <int B, typename T, typename VT>
__global__ __launch_bounds__(kBlock) void k_ilu_backward(int32_t cnt, const int32_t* __restrict__ rows,
                                                         const int32_t* __restrict__ rowptr,
                                                         const int32_t* __restrict__ col,
                                                         const int32_t* __restrict__ diag,
                                                         const VT* __restrict__ val, T* __restrict__ y) {
  const int32_t t = (int32_t)(blockIdx.x * kBlock + threadIdx.x);
  if (t >= cnt) return;
  const int32_t i = rows[t];
  const int32_t di = diag[i];
  T yi[B];
#pragma unroll
  for (int r = 0; r < B; ++r) yi[r] = y[(int64_t)i * B + r];
  for (int32_t pos = rowptr[i + 1] - 1; pos > di; --pos) {
    T yc[B];
#pragma unroll
    for (int c = 0; c < B; ++c) yc[c] = y[(int64_t)col[pos] * B + c];
    block_subtract<B>(yi, val + (int64_t)pos * B * B, yc);
  }
  const VT* D = val + (int64_t)di * B * B;
#pragma unroll
  for (int r = 0; r < B; ++r) {
    T a = (T)0;
#pragma unroll
    for (int c = 0; c < B; ++c) a = add_rn(a, mul_rn((T)D[r * B + c], yi[c]));
    y[(int64_t)i * B + r] = a;
  }
}

template <int B>
int factor_levels(gh_ctx* ctx, gh_ilu* I) {
  for (size_t L = 0; L + 1 < I->lptr_f.size(); ++L) {
    const int32_t a = I->lptr_f[L], cnt = I->lptr_f[L + 1] - a;
    hipLaunchKernelGGL(k_ilu_factor<B>, dim3(grid_for(cnt)), dim3(kBlock), 0, ctx->stream, cnt,
                       I->rows_f + a, I->rowptr, I->col, I->diag, I->val, I->d_flag);
  }
  GH_HIP(hipGetLastError());
  return GH_OK;
}

template <int B, typename T, typename VT>
int solve_levels(gh_ctx* ctx, const gh_ilu* I, const VT* val, T* y) {
  for (size_t L = 0; L + 1 < I->lptr_f.size(); ++L) {
    const int32_t a = I->lptr_f[L], cnt = I->lptr_f[L + 1] - a;
    hipLaunchKernelGGL((k_ilu_forward<B, T, VT>), dim3(grid_for(cnt)), dim3(kBlock), 0,
                       ctx->stream, cnt, I->rows_f + a, I->rowptr, I->col, I->diag, val, y);
  }
  for (size_t L = 0; L + 1 < I->lptr_b.size(); ++L) {
    const int32_t a = I->lptr_b[L], cnt = I->lptr_b[L + 1] - a;
    hipLaunchKernelGGL((k_ilu_backward<B, T, VT>), dim3(grid_for(cnt)), dim3(kBlock), 0,
                       ctx->stream, cnt, I->rows_b + a, I->rowptr, I->col, I->diag, val, y);
  }
  GH_HIP(hipGetLastError());
  return GH_OK;
}

template <typename T>
int dupload(gh_ctx* ctx, T** d, const std::vector<T>& h) {
  *d = nullptr;
  if (hipMalloc((void**)d, std::max<size_t>(h.size(), 1) * sizeof(T)) != hipSuccess)
    return fail(GH_ENOMEM, "gh_ilu_create: allocation failed");
  if (!h.empty())
    GH_HIP(hipMemcpyAsync(*d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, ctx->stream));
  return GH_OK;
}

// rows grouped by level: level(i) = 1 + the largest level of the rows its
// entries in [lo(i), hi(i)) reference (0 without any); ascending row order
// inside a level
void schedule(int64_t n, const std::vector<int32_t>& order_rows,
              const std::function<void(int32_t, int32_t&, int32_t&)>& range,
              const std::vector<int32_t>& col, std::vector<int32_t>& rows,
              std::vector<int32_t>& lptr) {
  std::vector<int32_t> lev(n, 0);
  int32_t nl = 0;
  for (int32_t i : order_rows) {
    int32_t lo, hi, l = 0;
    range(i, lo, hi);
    for (int32_t pos = lo; pos < hi; ++pos) l = std::max(l, lev[col[pos]] + 1);
    lev[i] = l;
    nl = std::max(nl, l + 1);
  }
  lptr.assign(nl + 1, 0);
  for (int64_t i = 0; i < n; ++i) lptr[lev[i] + 1]++;
  for (int32_t L = 0; L < nl; ++L) lptr[L + 1] += lptr[L];
  rows.resize(n);
  std::vector<int32_t> fill(lptr.begin(), lptr.end() - 1);
  for (int64_t i = 0; i < n; ++i) rows[fill[lev[i]]++] = (int32_t)i;
}

}  // namespace

namespace gh {
int ilu_solve(gh_ctx* ctx, const gh_ilu* I, int prec, void* x);
}

extern "C" {

int gh_ilu_create(gh_ctx* ctx, const gh_csr* A, const int32_t* perm, gh_ilu** out) {
  GH_ARG(ctx && A && out, "gh_ilu_create: NULL argument");
  GH_ARG(A->prec == GH_F64, "gh_ilu_create: needs an FP64 matrix");
  GH_ARG(A->ncomp >= 1 && A->ncomp <= 4, "gh_ilu_create: block size must be 1..4");
  GH_ARG(A->h_val_ok, "gh_ilu_create: the matrix values were assembled on the device");
  *out = nullptr;
  const int64_t n = A->n, nnz = A->nnz;
  const int B = A->ncomp, BB = B * B;
  // ConstructStructure
  std::vector<int32_t> p(n), q(n, -1);
  for (int64_t i = 0; i < n; ++i) {
    p[i] = perm ? perm[i] : (int32_t)i;
    if (p[i] < 0 || p[i] >= n || q[p[i]] >= 0)
      return fail(GH_EINVAL, "gh_ilu_create: the order is not a permutation");
    q[p[i]] = (int32_t)i;
  }
  const std::vector<int32_t>& arp = A->h_rowptr;
  const std::vector<int32_t>& acol = A->h_col;
  std::vector<int32_t> rowptr(n + 1, 0), col(nnz), diag(n, -1);
  for (int64_t i = 0; i < n; ++i) {
    const int32_t pi = p[i];
    const int32_t s = arp[pi], ni = arp[pi + 1] - arp[pi];
    rowptr[i + 1] = rowptr[i] + ni;
    for (int32_t c = 0; c < ni; ++c) col[rowptr[i] + c] = q[acol[s + c]];
    std::sort(col.begin() + rowptr[i], col.begin() + rowptr[i + 1]);
    for (int32_t pos = rowptr[i]; pos < rowptr[i + 1]; ++pos)
      if (col[pos] == i) diag[i] = pos;
    if (diag[i] < 0) return fail(GH_EINVAL, "gh_ilu_create: row %lld has no diagonal", (long long)i);
  }
  // copy_entries: block (p[i], j) of A added into (i, q[j])
  std::vector<double> val((size_t)nnz * BB, 0.0);
  for (int64_t i = 0; i < n; ++i) {
    const int32_t pi = p[i];
    for (int32_t posA = arp[pi]; posA < arp[pi + 1]; ++posA) {
      const int32_t pj = q[acol[posA]];
      for (int32_t pos = rowptr[i]; pos < rowptr[i + 1]; ++pos)
        if (col[pos] == pj) {
          for (int e = 0; e < BB; ++e)
            val[(size_t)pos * BB + e] += A->h_val[(size_t)posA * BB + e];
          break;
        }
    }
  }
  // level schedules of the lower (forward, factorisation) and upper parts
  std::vector<int32_t> asc(n), desc(n);
  std::iota(asc.begin(), asc.end(), 0);
  std::iota(desc.rbegin(), desc.rend(), 0);
  gh_ilu* I = new gh_ilu();
  I->ctx = ctx;
  I->n = n;
  I->nnz = nnz;
  I->B = B;
  std::vector<int32_t> rows_f, rows_b;
  schedule(n, asc, [&](int32_t i, int32_t& lo, int32_t& hi) { lo = rowptr[i]; hi = diag[i]; }, col,
           rows_f, I->lptr_f);
  schedule(n, desc, [&](int32_t i, int32_t& lo, int32_t& hi) { lo = diag[i] + 1; hi = rowptr[i + 1]; },
           col, rows_b, I->lptr_b);
  ctx_retain(ctx);  // balanced by gh_ilu_destroy
  int rc = dupload(ctx, &I->p, p);
  if (!rc) rc = dupload(ctx, &I->q, q);
  if (!rc) rc = dupload(ctx, &I->rowptr, rowptr);
  if (!rc) rc = dupload(ctx, &I->col, col);
  if (!rc) rc = dupload(ctx, &I->diag, diag);
  if (!rc) rc = dupload(ctx, &I->val, val);
  if (!rc) rc = dupload(ctx, &I->rows_f, rows_f);
  if (!rc) rc = dupload(ctx, &I->rows_b, rows_b);
  if (!rc && (hipMalloc((void**)&I->val32, sizeof(float) * std::max<size_t>(val.size(), 1)) != hipSuccess ||
              hipMalloc(&I->y, sizeof(double) * std::max<int64_t>(n * B, 1)) != hipSuccess ||
              hipMalloc((void**)&I->d_flag, sizeof(int)) != hipSuccess))
    rc = fail(GH_ENOMEM, "gh_ilu_create: allocation failed");
  if (!rc && hipMemsetAsync(I->d_flag, 0, sizeof(int), ctx->stream) != hipSuccess)
    rc = fail(GH_EHIP, "gh_ilu_create: memset");
  if (!rc) {
    switch (B) {
      case 1: rc = factor_levels<1>(ctx, I); break;
      case 2: rc = factor_levels<2>(ctx, I); break;
      case 3: rc = factor_levels<3>(ctx, I); break;
      default: rc = factor_levels<4>(ctx, I); break;
    }
  }
  if (!rc && nnz > 0) {
    const int64_t m = nnz * BB;
    hipLaunchKernelGGL((k_convert<float, double>), dim3((unsigned)std::min<int64_t>((m + kBlock - 1) / kBlock, 8192)),
                       dim3(kBlock), 0, ctx->stream, m, I->val32, (const double*)I->val);
    if (hipGetLastError() != hipSuccess) rc = fail(GH_EHIP, "gh_ilu_create: convert");
  }
  int flag = 0;
  if (!rc && hipMemcpyAsync(&flag, I->d_flag, sizeof(int), hipMemcpyDeviceToHost, ctx->stream) != hipSuccess)
    rc = fail(GH_EHIP, "gh_ilu_create: flag");
  if (!rc && hipStreamSynchronize(ctx->stream) != hipSuccess) rc = fail(GH_EHIP, "gh_ilu_create: sync");
  if (!rc && flag) rc = fail(GH_ESINGULAR, "gh_ilu_create: zero pivot in a diagonal block");
  if (rc) {
    gh_ilu_destroy(I);
    return rc;
  }
  *out = I;
  return GH_OK;
}

int gh_ilu_destroy(gh_ilu* I) {
  if (!I) return GH_OK;
  gh_ctx* ctx = I->ctx;
  if (ctx && ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  for (void* ptr : {(void*)I->p, (void*)I->q, (void*)I->rowptr, (void*)I->col, (void*)I->diag,
                    (void*)I->val, (void*)I->val32, I->y, (void*)I->rows_f, (void*)I->rows_b,
                    (void*)I->d_flag})
    if (ptr) (void)hipFree(ptr);
  delete I;
  if (ctx) ctx_release(ctx);
  return GH_OK;
}

int gh_ilu_levels(const gh_ilu* I, int32_t* forward, int32_t* backward) {
  GH_ARG(I && forward && backward, "gh_ilu_levels: NULL argument");
  *forward = (int32_t)I->lptr_f.size() - 1;
  *backward = (int32_t)I->lptr_b.size() - 1;
  return GH_OK;
}

int gh_ilu_download(gh_ctx* ctx, const gh_ilu* I, int32_t* rowptr, int32_t* col, int32_t* diag,
                    double* val) {
  GH_ARG(ctx && I, "gh_ilu_download: NULL argument");
  const int BB = I->B * I->B;
  if (rowptr) GH_HIP(hipMemcpy(rowptr, I->rowptr, sizeof(int32_t) * (I->n + 1), hipMemcpyDeviceToHost));
  if (col && I->nnz) GH_HIP(hipMemcpy(col, I->col, sizeof(int32_t) * I->nnz, hipMemcpyDeviceToHost));
  if (diag && I->n) GH_HIP(hipMemcpy(diag, I->diag, sizeof(int32_t) * I->n, hipMemcpyDeviceToHost));
  if (val && I->nnz)
    GH_HIP(hipMemcpy(val, I->val, sizeof(double) * I->nnz * BB, hipMemcpyDeviceToHost));
  return GH_OK;
}

int gh_ilu_solve(gh_ctx* ctx, const gh_ilu* I, gh_vec* x) {
  GH_ARG(ctx && I && x, "gh_ilu_solve: NULL argument");
  GH_ARG(x->n == I->n && x->ncomp == I->B, "gh_ilu_solve: vector shape does not match");
  return ilu_solve(ctx, I, x->prec, x->d);
}

}  // extern "C"

namespace gh {

// SparseBlockIlu::solve (:226-236), in place on x (FP64 or FP32)
int ilu_solve(gh_ctx* ctx, const gh_ilu* I, int prec, void* x) {
  const int64_t n = I->n;
  const int B = I->B;
  if (n == 0) return GH_OK;
  const unsigned g = grid_for(n * B);
  if (prec == GH_F64) {
    double* y = (double*)I->y;
    hipLaunchKernelGGL(k_ilu_gather<double>, dim3(g), dim3(kBlock), 0, ctx->stream, n, B, I->p,
                       (const double*)x, y);
    switch (B) {
      case 1: GH_TRY((solve_levels<1, double, double>(ctx, I, I->val, y))); break;
      case 2: GH_TRY((solve_levels<2, double, double>(ctx, I, I->val, y))); break;
      case 3: GH_TRY((solve_levels<3, double, double>(ctx, I, I->val, y))); break;
      default: GH_TRY((solve_levels<4, double, double>(ctx, I, I->val, y))); break;
    }
    hipLaunchKernelGGL(k_ilu_gather<double>, dim3(g), dim3(kBlock), 0, ctx->stream, n, B, I->q,
                       (const double*)y, (double*)x);
  } else {
    float* y = (float*)I->y;
    hipLaunchKernelGGL(k_ilu_gather<float>, dim3(g), dim3(kBlock), 0, ctx->stream, n, B, I->p,
                       (const float*)x, y);
    switch (B) {
      case 1: GH_TRY((solve_levels<1, float, float>(ctx, I, I->val32, y))); break;
      case 2: GH_TRY((solve_levels<2, float, float>(ctx, I, I->val32, y))); break;
      case 3: GH_TRY((solve_levels<3, float, float>(ctx, I, I->val32, y))); break;
      default: GH_TRY((solve_levels<4, float, float>(ctx, I, I->val32, y))); break;
    }
    hipLaunchKernelGGL(k_ilu_gather<float>, dim3(g), dim3(kBlock), 0, ctx->stream, n, B, I->q,
                       (const float*)y, (float*)x);
  }
  GH_HIP(hipGetLastError());
  return GH_OK;
}

}  // namespace gh
