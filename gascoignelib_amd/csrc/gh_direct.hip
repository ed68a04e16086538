// Direct solve of the coarsest multigrid levels: the dense inverse of the
// level matrix, factorised on the device (LU with partial pivoting and the
// inverse from it, rocSOLVER getrf / getri) and applied by k_dense_solve.
// The reference solves its coarse level with UMFPACK when it is built with it
// (src/LinAlg/umfilu.cc, StdSolver::DirectSolver) and otherwise with
// iterexact smoothing sweeps; any exact solve gives the same V-cycle up to
// rounding.
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>

#include "gh_kernels.h"

using namespace gh;

namespace {

// dense[i * m + j] += A(i, j) (row-major), node blocks expanded
__global__ __launch_bounds__(kBlock) void k_scatter_dense(int64_t n, int B, int64_t m,
                                                          const int32_t* __restrict__ rowptr,
                                                          const int32_t* __restrict__ col,
                                                          const double* __restrict__ val,
                                                          double* __restrict__ dense) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const int BB = B * B;
  for (int32_t pos = rowptr[i]; pos < rowptr[i + 1]; ++pos) {
    const int64_t j = col[pos];
    for (int r = 0; r < B; ++r)
      for (int c = 0; c < B; ++c)
        dense[(i * B + r) * m + j * B + c] += val[(int64_t)pos * BB + r * B + c];
  }
}

}  // namespace

namespace gh {

void direct_ctx_free(gh_ctx* ctx) {
  if (ctx && ctx->blas) {
    (void)rocblas_destroy_handle((rocblas_handle)ctx->blas);
    ctx->blas = nullptr;
  }
}

int dense_inverse_device(gh_ctx* ctx, const gh_csr* A, double* inv) {
  if (A->prec != GH_F64 || !A->val || !A->rowptr)
    return fail(GH_EINVAL, "direct coarse level: needs the FP64 device matrix");
  const int B = A->ncomp;
  const int64_t m = A->n * B;
  if (m == 0) return GH_OK;
  if (m > 46340) return fail(GH_EINVAL, "direct coarse level: %lld unknowns is too many for a "
                                        "dense inverse", (long long)m);
  if (!ctx->blas) {
    rocblas_handle h = nullptr;
    if (rocblas_create_handle(&h) != rocblas_status_success)
      return fail(GH_EHIP, "direct coarse level: rocblas_create_handle failed");
    ctx->blas = h;
  }
  rocblas_handle h = (rocblas_handle)ctx->blas;
  if (rocblas_set_stream(h, ctx->stream) != rocblas_status_success)
    return fail(GH_EHIP, "direct coarse level: rocblas_set_stream failed");
  rocblas_int* ipiv = nullptr;
  rocblas_int* info = nullptr;
  if (hipMalloc((void**)&ipiv, sizeof(rocblas_int) * m) != hipSuccess ||
      hipMalloc((void**)&info, sizeof(rocblas_int) * 2) != hipSuccess) {
    if (ipiv) (void)hipFree(ipiv);
    return fail(GH_ENOMEM, "direct coarse level: allocation failed");
  }
  int rc = GH_OK;
  // the row-major scatter is A^T in column-major order; the column-major
  // inverse of A^T read row-major is A^-1
  hipLaunchKernelGGL(k_fill0<double>, dim3((unsigned)std::min<int64_t>((m * m + kBlock - 1) / kBlock, 8192)),
                     dim3(kBlock), 0, ctx->stream, m * m, inv);
  hipLaunchKernelGGL(k_scatter_dense, dim3(grid_for(A->n)), dim3(kBlock), 0, ctx->stream, A->n,
                     B, m, A->rowptr, A->col, (const double*)A->val, inv);
  if (hipGetLastError() != hipSuccess) rc = fail(GH_EHIP, "direct coarse level: scatter");
  if (!rc && rocsolver_dgetrf(h, (rocblas_int)m, (rocblas_int)m, inv, (rocblas_int)m, ipiv,
                              info) != rocblas_status_success)
    rc = fail(GH_EHIP, "direct coarse level: rocsolver_dgetrf failed");
  if (!rc && rocsolver_dgetri(h, (rocblas_int)m, inv, (rocblas_int)m, ipiv, info + 1) !=
                 rocblas_status_success)
    rc = fail(GH_EHIP, "direct coarse level: rocsolver_dgetri failed");
  rocblas_int hinfo[2] = {0, 0};
  if (!rc && hipMemcpyAsync(hinfo, info, sizeof(hinfo), hipMemcpyDeviceToHost, ctx->stream) !=
                 hipSuccess)
    rc = fail(GH_EHIP, "direct coarse level: info copy");
  if (!rc && hipStreamSynchronize(ctx->stream) != hipSuccess)
    rc = fail(GH_EHIP, "direct coarse level: sync");
  if (!rc && (hinfo[0] != 0 || hinfo[1] != 0))
    rc = fail(GH_ESINGULAR, "direct coarse level: singular matrix (pivot %d)", (int)hinfo[0]);
  (void)hipFree(ipiv);
  (void)hipFree(info);
  return rc;
}

}  // namespace gh
