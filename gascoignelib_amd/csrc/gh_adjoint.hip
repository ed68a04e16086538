// Adjoint pieces of the gascoignelib_amd C ABI: the transposed operators the
// adjoint multigrid solve runs on, and the per-cell coefficient gradient
//   dJ/dk_c = -z^T (dA/dk_c) u
// of a functional J(u(k)) of the solution of A(k) u = f, with A^T z = dJ/du.
//
// The reference has no adjoint solve (StdSolver's dual-matrix path,
// src/Solver/stdsolver.cc:1068-1090, 1815-1834, stops at assembly), so this
// file has no counterpart there.  The coefficient enters the assembly cell by
// cell (A = sum_c k_c K^c with Dirichlet rows and columns eliminated), so
// dA/dk_c is the unit-coefficient element matrix K^c on the free nodes.
#include <algorithm>
#include <cmath>
#include <vector>

#include "gh_internal.h"

using namespace gh;

namespace gh {

void csr_transpose_host(int64_t n, int B, const int32_t* rowptr, const int32_t* col,
                        const double* val, std::vector<int32_t>& rpT, std::vector<int32_t>& clT,
                        std::vector<double>& vlT) {
  const int BB = B * B;
  const int64_t nnz = rowptr[n];
  rpT.assign(n + 1, 0);
  for (int64_t q = 0; q < nnz; ++q) rpT[col[q] + 1]++;
  for (int64_t i = 0; i < n; ++i) rpT[i + 1] += rpT[i];
  clT.resize(nnz);
  vlT.resize(nnz * BB);
  std::vector<int32_t> fill(rpT.begin(), rpT.end() - 1);
  // rows visited in ascending order: each transposed row lists its columns ascending
  for (int64_t i = 0; i < n; ++i)
    for (int32_t q = rowptr[i]; q < rowptr[i + 1]; ++q) {
      const int32_t t = fill[col[q]]++;
      clT[t] = (int32_t)i;
      for (int a = 0; a < B; ++a)
        for (int b = 0; b < B; ++b) vlT[(int64_t)t * BB + a * B + b] = val[(int64_t)q * BB + b * B + a];
    }
}

}  // namespace gh

namespace {

// 1D Q1 element on [0, h]: stiffness and mass
__device__ __forceinline__ double s1(int p, int q, double h) { return (p == q ? 1.0 : -1.0) / h; }
__device__ __forceinline__ double m1(int p, int q, double h) { return h * (p == q ? 2.0 : 1.0) / 6.0; }

// Corner a of a hex sits at (a&1 ^ (a>>1&1), a>>1&1, a>>2): 0 (0,0,0), 1 (1,0,0),
// 2 (1,1,0), 3 (0,1,0), then the same on the top face (Gascoigne's order).
__device__ __forceinline__ int cx(int a) { return (a & 1) ^ ((a >> 1) & 1); }
__device__ __forceinline__ int cy(int a) { return (a >> 1) & 1; }
__device__ __forceinline__ int cz(int a) { return (a >> 2) & 1; }

// One thread per cell: g[c] = -sum_{a,b} z(v_a) K^c_ab u(v_b)
template <int DIM>
__global__ __launch_bounds__(kBlock) void k_dJdk(int64_t ncells, const int32_t* __restrict__ cells,
                                                 const double* __restrict__ h,
                                                 const double* __restrict__ z,
                                                 const double* __restrict__ u,
                                                 double* __restrict__ g) {
  const int64_t c = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (c >= ncells) return;
  constexpr int NPC = DIM == 3 ? 8 : 4;
  const int32_t* v = cells + c * NPC;
  double zv[NPC], uv[NPC];
#pragma unroll
  for (int a = 0; a < NPC; ++a) {
    zv[a] = z[v[a]];
    uv[a] = u[v[a]];
  }
  const double hx = h[3 * c], hy = h[3 * c + 1], hz = h[3 * c + 2];
  double s = 0.0;
#pragma unroll
  for (int a = 0; a < NPC; ++a) {
    double ka = 0.0;
#pragma unroll
    for (int b = 0; b < NPC; ++b) {
      double e;
      if (DIM == 3)
        e = s1(cx(a), cx(b), hx) * m1(cy(a), cy(b), hy) * m1(cz(a), cz(b), hz) +
            m1(cx(a), cx(b), hx) * s1(cy(a), cy(b), hy) * m1(cz(a), cz(b), hz) +
            m1(cx(a), cx(b), hx) * m1(cy(a), cy(b), hy) * s1(cz(a), cz(b), hz);
      else
        e = s1(cx(a), cx(b), hx) * m1(cy(a), cy(b), hy) +
            m1(cx(a), cx(b), hx) * s1(cy(a), cy(b), hy);
      ka += e * uv[b];
    }
    s += zv[a] * ka;
  }
  g[c] = -s;
}

}  // namespace

extern "C" {

int gh_cells_create(gh_ctx* ctx, int dim, int64_t n, const double* xyz, int64_t ncells,
                    const int32_t* cells, gh_cells** out) {
  GH_ARG(ctx && out && xyz && (cells || ncells == 0) && n >= 0 && ncells >= 0,
         "gh_cells_create: bad arguments");
  GH_ARG(dim == 2 || dim == 3, "gh_cells_create: dim must be 2 or 3");
  GH_ARG(ncells < (int64_t(1) << 28) && n < (int64_t(1) << 31),
         "gh_cells_create: too many cells or nodes");
  const int npc = dim == 3 ? 8 : 4;
  static const int corner[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                                   {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
  std::vector<double> h(3 * ncells, 1.0);
  for (int64_t c = 0; c < ncells; ++c) {
    const int32_t* v = cells + c * npc;
    for (int m = 0; m < npc; ++m)
      if (v[m] < 0 || v[m] >= n) return fail(GH_EINVAL, "gh_cells_create: node out of range in cell %lld", (long long)c);
    const double* x0 = xyz + 3 * (int64_t)v[0];
    h[3 * c] = xyz[3 * (int64_t)v[1]] - x0[0];
    h[3 * c + 1] = xyz[3 * (int64_t)v[3] + 1] - x0[1];
    if (dim == 3) h[3 * c + 2] = xyz[3 * (int64_t)v[4] + 2] - x0[2];
    for (int m = 0; m < npc; ++m) {
      const double* xm = xyz + 3 * (int64_t)v[m];
      for (int d = 0; d < dim; ++d) {
        const double e = x0[d] + corner[m][d] * h[3 * c + d];
        if (std::fabs(xm[d] - e) > 1e-12 * (std::fabs(e) + 1.0))
          return fail(GH_EINVAL, "gh_cells_create: cell %lld is not an axis-aligned box", (long long)c);
      }
    }
    if (h[3 * c] <= 0 || h[3 * c + 1] <= 0 || h[3 * c + 2] <= 0)
      return fail(GH_EINVAL, "gh_cells_create: cell %lld has inverted orientation", (long long)c);
  }
  // node -> (cell, corner), ascending cell per node (counting sort)
  std::vector<int32_t> nptr(n + 1, 0), nlist((size_t)ncells * npc);
  for (int64_t q = 0; q < ncells * npc; ++q) nptr[cells[q] + 1]++;
  for (int64_t i = 0; i < n; ++i) nptr[i + 1] += nptr[i];
  {
    std::vector<int32_t> fill(nptr.begin(), nptr.end() - 1);
    for (int64_t c = 0; c < ncells; ++c)
      for (int m = 0; m < npc; ++m) nlist[fill[cells[c * npc + m]]++] = (int32_t)(c * 8 + m);
  }
  gh_cells* C = new gh_cells();
  C->ctx = ctx;
  ctx_retain(ctx);  // balanced by gh_cells_destroy
  C->dim = dim;
  C->n = n;
  C->ncells = ncells;
  auto cleanup = [&](int rc) {
    gh_cells_destroy(C);
    return rc;
  };
  if (ncells > 0) {
    if (hipMalloc(&C->cells, sizeof(int32_t) * ncells * npc) != hipSuccess ||
        hipMalloc(&C->h, sizeof(double) * 3 * ncells) != hipSuccess)
      return cleanup(fail(GH_ENOMEM, "gh_cells_create: allocation failed"));
    if (hipMemcpy(C->cells, cells, sizeof(int32_t) * ncells * npc, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(C->h, h.data(), sizeof(double) * 3 * ncells, hipMemcpyHostToDevice) != hipSuccess)
      return cleanup(fail(GH_EHIP, "gh_cells_create: upload failed"));
  }
  if (hipMalloc(&C->nptr, sizeof(int32_t) * (n + 1)) != hipSuccess ||
      hipMalloc(&C->nlist, sizeof(int32_t) * std::max<int64_t>(ncells * npc, 1)) != hipSuccess)
    return cleanup(fail(GH_ENOMEM, "gh_cells_create: allocation failed"));
  if (hipMemcpy(C->nptr, nptr.data(), sizeof(int32_t) * (n + 1), hipMemcpyHostToDevice) != hipSuccess ||
      (ncells > 0 && hipMemcpy(C->nlist, nlist.data(), sizeof(int32_t) * ncells * npc,
                               hipMemcpyHostToDevice) != hipSuccess))
    return cleanup(fail(GH_EHIP, "gh_cells_create: upload failed"));
  *out = C;
  return GH_OK;
}

int gh_cells_destroy(gh_cells* C) {
  if (!C) return GH_OK;
  if (C->cells) (void)hipFree(C->cells);
  if (C->h) (void)hipFree(C->h);
  if (C->nptr) (void)hipFree(C->nptr);
  if (C->nlist) (void)hipFree(C->nlist);
  gh_ctx* ctx = C->ctx;
  delete C;
  if (ctx) ctx_release(ctx);
  return GH_OK;
}

int gh_dJdk_cells(gh_ctx* ctx, const gh_cells* C, const gh_vec* z, const gh_vec* u, gh_vec* g) {
  GH_ARG(ctx && C && z && u && g, "gh_dJdk_cells: NULL argument");
  GH_ARG(z->prec == GH_F64 && u->prec == GH_F64 && g->prec == GH_F64,
         "gh_dJdk_cells: vectors must be FP64");
  GH_ARG(z->ncomp == 1 && u->ncomp == 1 && z->n == C->n && u->n == C->n,
         "gh_dJdk_cells: z and u must be scalar nodal vectors of the mesh");
  GH_ARG(g->n * g->ncomp == C->ncells, "gh_dJdk_cells: g must have one entry per cell");
  if (C->ncells == 0) return GH_OK;
  const dim3 grid((unsigned)((C->ncells + kBlock - 1) / kBlock));
  if (C->dim == 3)
    hipLaunchKernelGGL(k_dJdk<3>, grid, dim3(kBlock), 0, ctx->stream, C->ncells, C->cells, C->h,
                       (const double*)z->d, (const double*)u->d, (double*)g->d);
  else
    hipLaunchKernelGGL(k_dJdk<2>, grid, dim3(kBlock), 0, ctx->stream, C->ncells, C->cells, C->h,
                       (const double*)z->d, (const double*)u->d, (double*)g->d);
  GH_HIP(hipGetLastError());
  return GH_OK;
}

}  // extern "C"
