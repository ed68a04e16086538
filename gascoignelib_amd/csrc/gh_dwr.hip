// Dual-weighted residual indicators of a Q1 solution on the device (the
// "DWR" estimator StdLoop::Estimator would call; src/Solver/stdloop.cc:221-273
// holds only stubs that abort).
//
// For every cell K of a patch P (the 2^d children of one coarser cell, 3^d
// nodes, an axis-aligned box with midpoint middle nodes):
//
//   eta_K = int_K f w - k_K int_K grad u_h . grad w,     w = I2 z_h - z_h,
//
// u_h, z_h the Q1 functions of the nodal primal and adjoint vectors, I2 z_h
// the Q2 Lagrange interpolant of z on the patch, f the Q2 interpolant of the
// nodal data f on the patch, k_K a constant per cell.  Sum_K eta_K estimates
// J(u) - J(u_h) for the functional whose adjoint z is (Becker & Rannacher,
// Acta Numerica 10 (2001), sections 3-4).  One thread per (patch, cell)
// integrates with the 3^d-point Gauss rule on the cell, exact for these
// integrands (degree <= 4 per direction).
#include <climits>
#include <cmath>
#include <vector>

#include "gh_kernels.h"

using namespace gh;

namespace {

template <int DIM>
__global__ __launch_bounds__(kBlock) void k_dwr(int64_t npatches, const int32_t* __restrict__ nodes,
                                                const double* __restrict__ hp,
                                                const double* __restrict__ u,
                                                const double* __restrict__ z,
                                                const double* __restrict__ f,
                                                const double* __restrict__ kc,
                                                double* __restrict__ eta) {
  constexpr int NC = 1 << DIM;                       // cells per patch
  constexpr int NN = DIM == 3 ? 27 : 9;              // nodes per patch
  const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (t >= npatches * NC) return;
  const int64_t p = t / NC;
  const int c = (int)(t - p * NC);
  int sub[DIM];
#pragma unroll
  for (int d = 0; d < DIM; ++d) sub[d] = (c >> d) & 1;
  double h[DIM];
  double vol = 1.0;
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    h[d] = hp[p * DIM + d];
    vol *= h[d];
  }
  const int32_t* nd = nodes + p * NN;
  double zp[NN], fp[NN];
#pragma unroll
  for (int j = 0; j < NN; ++j) {
    zp[j] = z[nd[j]];
    fp[j] = f[nd[j]];
  }
  // the cell's corners: patch-local index sum_d (sub_d + a_d) 3^d
  double uc[NC], zc[NC];
#pragma unroll
  for (int a = 0; a < NC; ++a) {
    int j = 0, s = 1;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      j += (sub[d] + ((a >> d) & 1)) * s;
      s *= 3;
    }
    uc[a] = u[nd[j]];
    zc[a] = zp[j];
  }
  const double k = kc ? kc[t] : 1.0;
  const double gq[3] = {0.5 - 0.5 * 0.7745966692414834, 0.5, 0.5 + 0.5 * 0.7745966692414834};
  const double gw[3] = {5.0 / 18.0, 8.0 / 18.0, 5.0 / 18.0};
  double acc = 0.0;
  int q[DIM];
  for (int qi = 0; qi < (DIM == 3 ? 27 : 9); ++qi) {
    q[0] = qi % 3;
    q[1] = (qi / 3) % 3;
    if (DIM == 3) q[DIM - 1] = qi / 9;
    double wq = 1.0;
    double L[DIM][3], dL[DIM][3], N[DIM][2], dN[DIM][2];
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      const double xi = gq[q[d]];        // in the cell
      const double s = sub[d] + xi;      // in the patch, nodes at 0, 1, 2
      wq *= gw[q[d]];
      L[d][0] = 0.5 * (s - 1.0) * (s - 2.0);
      L[d][1] = s * (2.0 - s);
      L[d][2] = 0.5 * s * (s - 1.0);
      dL[d][0] = s - 1.5;
      dL[d][1] = 2.0 - 2.0 * s;
      dL[d][2] = s - 0.5;
      N[d][0] = 1.0 - xi;
      N[d][1] = xi;
      dN[d][0] = -1.0;
      dN[d][1] = 1.0;
    }
    // Q2 interpolants of z and f: value and (z only) gradient in patch units
    double z2 = 0.0, f2 = 0.0, gz2[DIM] = {};
#pragma unroll
    for (int j = 0; j < NN; ++j) {
      int a[DIM];
      a[0] = j % 3;
      a[1] = (j / 3) % 3;
      if (DIM == 3) a[DIM - 1] = j / 9;
      double v = 1.0;
#pragma unroll
      for (int d = 0; d < DIM; ++d) v *= L[d][a[d]];
      z2 += v * zp[j];
      f2 += v * fp[j];
#pragma unroll
      for (int g = 0; g < DIM; ++g) {
        double dv = 1.0;
#pragma unroll
        for (int d = 0; d < DIM; ++d) dv *= (d == g) ? dL[d][a[d]] : L[d][a[d]];
        gz2[g] += dv * zp[j];
      }
    }
    // Q1 on the cell: z value and gradient, u gradient
    double z1 = 0.0, gz1[DIM] = {}, gu[DIM] = {};
#pragma unroll
    for (int a = 0; a < NC; ++a) {
      double v = 1.0;
#pragma unroll
      for (int d = 0; d < DIM; ++d) v *= N[d][(a >> d) & 1];
      z1 += v * zc[a];
#pragma unroll
      for (int g = 0; g < DIM; ++g) {
        double dv = 1.0;
#pragma unroll
        for (int d = 0; d < DIM; ++d) dv *= (d == g) ? dN[d][(a >> d) & 1] : N[d][(a >> d) & 1];
        gz1[g] += dv * zc[a];
        gu[g] += dv * uc[a];
      }
    }
    double grad = 0.0;
#pragma unroll
    for (int g = 0; g < DIM; ++g) grad += gu[g] * (gz2[g] - gz1[g]) / (h[g] * h[g]);
    acc += wq * (f2 * (z2 - z1) - k * grad);
  }
  eta[t] = acc * vol;
}

}  // namespace

extern "C" {

int gh_patches_create(gh_ctx* ctx, int dim, int64_t nnodes, const double* xyz, int64_t npatches,
                      const int32_t* nodes, gh_patches** out) {
  GH_ARG(ctx && out && (dim == 2 || dim == 3) && nnodes >= 0 && npatches >= 0 &&
             (npatches == 0 || (xyz && nodes)),
         "gh_patches_create: bad arguments");
  *out = nullptr;
  const int nn = dim == 3 ? 27 : 9;
  std::vector<double> h((size_t)npatches * dim);
  for (int64_t p = 0; p < npatches; ++p) {
    const int32_t* nd = nodes + p * nn;
    for (int j = 0; j < nn; ++j)
      if (nd[j] < 0 || nd[j] >= nnodes) return fail(GH_EINVAL, "gh_patches_create: node out of range");
    // an axis-aligned box: node (a_0, a_1, a_2) at x0 + a_d h_d along axis d
    const double* x0 = xyz + (int64_t)nd[0] * 3;
    const double* x1 = xyz + (int64_t)nd[nn - 1] * 3;
    for (int d = 0; d < dim; ++d) {
      h[p * dim + d] = 0.5 * (x1[d] - x0[d]);
      if (!(h[p * dim + d] > 0.0))
        return fail(GH_EINVAL, "gh_patches_create: patch %lld is not a box in lexicographic order",
                    (long long)p);
    }
    for (int j = 0; j < nn; ++j) {
      const int a[3] = {j % 3, (j / 3) % 3, j / 9};
      const double* xj = xyz + (int64_t)nd[j] * 3;
      for (int d = 0; d < dim; ++d) {
        const double want = x0[d] + a[d] * h[p * dim + d];
        if (std::fabs(xj[d] - want) > 1e-12 * (std::fabs(x0[d]) + std::fabs(x1[d]) + 1.0))
          return fail(GH_EINVAL,
                      "gh_patches_create: patch %lld is not an axis-aligned box with midpoint "
                      "nodes", (long long)p);
      }
    }
  }
  // the steps from node 0 to its neighbours along x, y, z (nodes 1, 3, 9),
  // the patch element's box as the host producer measures it
  std::vector<double> hs((size_t)npatches * 3, 1.0);
  for (int64_t p = 0; p < npatches; ++p)
    for (int d = 0; d < dim; ++d) {
      const int32_t* nd = nodes + p * nn;
      hs[p * 3 + d] = xyz[(int64_t)nd[d == 0 ? 1 : (d == 1 ? 3 : 9)] * 3 + d] - xyz[(int64_t)nd[0] * 3 + d];
    }
  // node -> (patch, local node), ascending patch
  if (npatches > (int64_t)(INT_MAX / 32))
    return fail(GH_EINVAL, "gh_patches_create: too many patches");
  std::vector<int32_t> pptr(nnodes + 1, 0), plist((size_t)npatches * nn);
  for (int64_t k = 0; k < npatches * nn; ++k) pptr[nodes[k] + 1]++;
  for (int64_t i = 0; i < nnodes; ++i) pptr[i + 1] += pptr[i];
  {
    std::vector<int32_t> fill(pptr.begin(), pptr.end() - 1);
    for (int64_t p = 0; p < npatches; ++p)
      for (int a = 0; a < nn; ++a) plist[fill[nodes[p * nn + a]]++] = (int32_t)(p * 32 + a);
  }
  gh_patches* P = new gh_patches();
  P->ctx = ctx;
  P->dim = dim;
  P->nnodes = nnodes;
  P->npatches = npatches;
  ctx_retain(ctx);  // balanced by gh_patches_destroy
  auto up = [&](void** d, const void* src, size_t bytes) {
    if (hipMalloc(d, std::max<size_t>(bytes, 8)) != hipSuccess) return false;
    return bytes == 0 || hipMemcpy(*d, src, bytes, hipMemcpyHostToDevice) == hipSuccess;
  };
  if (!up((void**)&P->nodes, nodes, sizeof(int32_t) * npatches * nn) ||
      !up((void**)&P->h, h.data(), sizeof(double) * npatches * dim) ||
      !up((void**)&P->hs, hs.data(), sizeof(double) * npatches * 3) ||
      !up((void**)&P->pptr, pptr.data(), sizeof(int32_t) * (nnodes + 1)) ||
      !up((void**)&P->plist, plist.data(), sizeof(int32_t) * npatches * nn)) {
    gh_patches_destroy(P);
    return fail(GH_ENOMEM, "gh_patches_create: allocation or upload failed");
  }
  *out = P;
  return GH_OK;
}

int gh_patches_destroy(gh_patches* P) {
  if (!P) return GH_OK;
  gh_ctx* ctx = P->ctx;
  if (ctx && ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  for (void* d : {(void*)P->nodes, (void*)P->h, (void*)P->hs, (void*)P->pptr, (void*)P->plist})
    if (d) (void)hipFree(d);
  delete P;
  if (ctx) ctx_release(ctx);
  return GH_OK;
}

int gh_dwr_estimate(gh_ctx* ctx, const gh_patches* P, const gh_vec* u, const gh_vec* z,
                    const gh_vec* f, const gh_vec* k, gh_vec* eta) {
  GH_ARG(ctx && P && u && z && f && eta, "gh_dwr_estimate: NULL argument");
  const int nc = 1 << P->dim;
  for (const gh_vec* v : {u, z, f})
    GH_ARG(v->prec == GH_F64 && v->ncomp == 1 && v->n == P->nnodes,
           "gh_dwr_estimate: u, z and f must be scalar FP64 nodal vectors of the patches' mesh");
  GH_ARG(eta->prec == GH_F64 && eta->size == P->npatches * nc,
         "gh_dwr_estimate: eta must hold one FP64 value per cell (npatches * 2^dim)");
  GH_ARG(!k || (k->prec == GH_F64 && k->size == P->npatches * nc),
         "gh_dwr_estimate: k must hold one FP64 value per cell (npatches * 2^dim)");
  const int64_t nt = P->npatches * nc;
  if (nt == 0) return GH_OK;
  const double* kd = k ? (const double*)k->d : nullptr;
  if (P->dim == 3)
    hipLaunchKernelGGL(k_dwr<3>, dim3(grid_for(nt)), dim3(kBlock), 0, ctx->stream, P->npatches,
                       P->nodes, P->h, (const double*)u->d, (const double*)z->d,
                       (const double*)f->d, kd, (double*)eta->d);
  else
    hipLaunchKernelGGL(k_dwr<2>, dim3(grid_for(nt)), dim3(kBlock), 0, ctx->stream, P->npatches,
                       P->nodes, P->h, (const double*)u->d, (const double*)z->d,
                       (const double*)f->d, kd, (double*)eta->d);
  GH_HIP(hipGetLastError());
  return GH_OK;
}

}  // extern "C"
