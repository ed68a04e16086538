// Multigrid V/W/F cycle, left-preconditioned GMRES and the Newton/iterative
// refinement driver of the gascoignelib_amd C ABI.
//
// The control flow follows StdMultiLevelSolver::LinearMg / mgstep
// (src/Solver/stdmultilevelsolver.cc:411-531), GMRES::solve
// (src/LinAlg/gmres.cc:160-260) and StdMultiLevelSolver::newton
// (stdmultilevelsolver.cc:548-758); every vector operation runs on the device.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "gh_sell.h"

using namespace gh;

// ===================================================================== info
extern "C" {

void gh_cginfo_init(gh_cginfo* I, double tol, double globaltol, int maxiter) {
  std::memset(I, 0, sizeof(*I));
  I->tol = tol;
  I->globaltol = globaltol;
  I->maxiter = maxiter;
  I->breaktol = 1.e15;
  gh_cginfo_reset(I);
}

// CGInfo::reset (src/Common/cginfo.cc:193-199) -> ControlData::reset (:106-115)
void gh_cginfo_reset(gh_cginfo* I) {
  I->residual = 0.;
  I->correction = 0.;
  I->iteration = 0;
  I->status = GH_WAITING;
  I->previousresidual = 1.;
  I->previouscorrection = 1.;
  I->rate = 0.;
  I->lastrate = 0.;
}

// CGInfo::check (src/Common/cginfo.cc:202-256)
int gh_cginfo_check(gh_cginfo* I, double resi, double cori) {
  const double res = std::fabs(resi), cor = std::fabs(cori);
  I->residual = res;
  I->correction = cor;
  if (I->status != GH_RUNNING) {
    I->firstresidual = res;
    I->previousresidual = res;
    I->previouscorrection = cor;
    I->aimedresidual = res * I->tol;
    I->status = GH_RUNNING;
  } else {
    I->totaliter++;
    I->lastrate = res / I->previousresidual;
    I->rate = std::pow(I->residual / I->firstresidual, 1. / (double)std::max(1, I->iteration));
    I->previousresidual = res;
  }
  if (I->residual < std::max(I->aimedresidual, I->globaltol))
    I->status = GH_CONVERGED;
  else if (I->iteration >= I->maxiter)
    I->status = GH_TOO_MANY;
  else if (I->residual > I->breaktol)
    I->status = GH_EXPLODED;
  else if (I->residual > I->previousresidual)
    I->status = GH_DIVERGING;
  I->iteration++;
  if (I->status == GH_RUNNING) return 0;
  if (I->iteration < I->miniter) return 0;
  return 1;
}

// NLInfo (src/Common/nlinfo.cc:100-356)
void gh_nlinfo_init(gh_nlinfo* I, double tol, double globaltol, int maxiter) {
  std::memset(I, 0, sizeof(*I));
  I->tol = tol;
  I->globaltol = globaltol;
  I->maxiter = maxiter;
  I->breaktol = 1.e15;
  I->maxrelax = 11;
  I->rho = 0.3;
  I->linrho = 1.;
  I->maxresincrease = 1.e3;
  gh_nlinfo_reset(I);
}

void gh_nlinfo_reset(gh_nlinfo* I) {
  I->rate = 0.;
  I->residual = 0.;
  I->correction = 0.;
  I->iteration = 0;
  I->status = GH_NL_WAITING;
  I->previousresidual = 1.;
  I->previouscorrection = 1.;
  I->relax = 0;
  I->laststepbad = 0;
}

int gh_nlinfo_check(gh_nlinfo* I, int iter, double resi, double cori) {
  const double res = std::fabs(resi), cor = std::fabs(cori);
  int newiteration = 0;
  if (I->status == GH_NL_DIVERGED || I->status == GH_NL_STAGNATION) return 1;
  if (I->status == GH_NL_WAITING) {
    newiteration = 1;
  } else if (iter > I->iteration) {
    newiteration = 1;
    I->iteration++;
  }
  I->status = GH_NL_RUNNING;
  int thisstepbad = 0;
  if (!I->iteration) {
    I->residual = res;
    I->previousresidual = res;
    I->correction = cor;
    I->firstresidual = res;
    I->aimedresidual = res * I->tol;
  } else {
    const double r0 = I->residual, c0 = I->correction;
    I->residual = res;
    I->correction = cor;
    I->rate = std::pow(I->residual / I->firstresidual, 1. / (double)std::max(1, I->iteration));
    I->lastrate = I->iteration > 1 ? I->residual / I->previousresidual : I->rate;
    if (I->residual > 2. * I->previousresidual) thisstepbad = 1;
    if (newiteration) {
      I->previousresidual = r0;
      I->previouscorrection = c0;
    }
  }
  if (thisstepbad && I->laststepbad)
    I->status = GH_NL_DIVERGED;
  else if (I->residual < std::max(I->aimedresidual, I->globaltol))
    I->status = GH_NL_CONVERGED;
  else if (I->iteration >= I->maxiter && I->rate < 1.)
    I->status = GH_NL_SLOW;
  else if (I->iteration >= I->maxiter)
    I->status = GH_NL_STAGNATION;
  else if (I->residual > I->breaktol)
    I->status = GH_NL_EXPLODED;
  else if (I->iteration && I->lastrate > I->maxresincrease)
    I->status = GH_NL_EXPLODED;
  I->laststepbad = thisstepbad;
  if (I->status == GH_NL_RUNNING) return 0;
  if (I->iteration < I->miniter) return 0;
  I->totaliter += I->iteration;
  return 1;
}

// NLInfo::check_damping (nlinfo.cc:223-238): 0 ok, 1 continue, 2 exploded
static int nl_check_damping(gh_nlinfo* I, int dampit, double res) {
  I->residual = std::fabs(res);
  I->relax = dampit;
  if (I->residual > 1.e20) return 2;
  if (!(I->residual <= 1.e20)) return 2;
  if (I->residual <= I->previousresidual) return 0;
  return 1;
}

void gh_mg_params_default(gh_mg_params* p) {
  p->mgtype = 'V';
  p->iterpre = 4;
  p->iterpost = 4;
  p->iterexact = 10;
  p->omega = 1.;
  p->mgomega = 1.;
  p->smoother = 0;
  p->ndirect = 1000;
  p->prec_maxiter = 1;
  p->prec_tol = 1.e-12;
  p->prec_globaltol = 1.e-12;
  p->prec = GH_F64;
  p->use_graph = 0;
  p->krylov = 0;
  p->adjoint = 0;
}

}  // extern "C"

// ================================================================ multigrid
namespace {

struct MgLevel {
  gh_csr* A64 = nullptr;   // FP64 matrix (the caller's, or an owned renumbered copy)
  gh_csr* A = nullptr;     // matrix of the V-cycle (A64 on the FP64 path)
  bool own_A = false;
  // finest level of a renumbered flexible-GMRES solver: FP64 copy of A64 whose
  // rows take their entries in ascending internal column order (shared
  // relative patterns), for the outer products and residuals, which reproduce
  // no reference sums
  gh_csr* A64s = nullptr;
  bool own_A64 = false, own_T = false, own_H = false, own_M = false;
  int32_t* d_perm = nullptr;  // internal position -> caller's node (nullptr: identity)
  gh_transfer* T = nullptr;
  gh_hn* H = nullptr;
  gh_mask* M = nullptr;
  int64_t n = 0, size = 0;
  int ncomp = 1;
  // restriction into level l-1 with its strong Dirichlet rows emptied
  int32_t* r_row = nullptr;
  uint32_t* r_col = nullptr;
  // prolongation from level l-1 with hanging / Dirichlet rows of this level emptied
  int32_t* p_row = nullptr;
  uint32_t* p_col = nullptr;
  Sell r_sell, p_sell;           // SELL-64 copies of the folded operators (ncomp == 1)
  bool fused_correction = true;  // mask covers every component
  bool mask_in_restrict = true;  // coarse Dirichlet rows emptied in r_row (set on level l)
  gh_vanka* vanka = nullptr;     // patch smoother (smoother 2), owned
  gh_ilu* ilu = nullptr;         // ILU(0) smoother (smoother 3), owned
  bool pre_fused = false;        // the first pre-smoothing sweep already ran (in the
                                 // restriction that produced b); consumed by mgstep
  void* u = nullptr;
  void* b = nullptr;
  void* v = nullptr;
  bool direct = false;
  double* Ainv = nullptr;
};

}  // namespace

struct gh_mg {
  gh_ctx* ctx = nullptr;
  int nlevels = 0;
  gh_mg_params P;
  int wprec = GH_F64;  // working precision of vectors
  std::vector<MgLevel> L;
  int64_t mgsteps = 0;
  // GMRES memory (FP64, finest level)
  std::vector<double*> mem;
  double* p = nullptr;
  double* r = nullptr;
  double* w = nullptr;
  // flexible GMRES: preconditioned directions z_j in the V-cycle's precision
  std::vector<void*> zmem;
  // solution update x += sum_j c_j basis_j: device pointer tables and the
  // coefficients, staged through pinned host memory
  const void** d_basis = nullptr;  // mem[0..vmax] then zmem[0..vmax]
  double* d_coef = nullptr;
  double* h_coef = nullptr;
  int basis_cap = 0;
  // fine-level staging for graph replay / precision change
  void* fine_b = nullptr;
  // renumbered solver: the caller's finest-level matrix handle and FP64
  // buffers holding caller vectors in the internal numbering
  bool ordered = false;
  const gh_csr* user_fine = nullptr;
  // the caller's finest-level Dirichlet mask and hanging-node handles (the
  // only ones gh_gmres_solve / gh_newton_solve accept besides NULL)
  const gh_mask* user_mask = nullptr;
  const gh_hn* user_hn = nullptr;
  double* ext_x = nullptr;
  double* ext_y = nullptr;
  // captured V-cycle (fast path from zero initial guess)
  hipGraphExec_t graph = nullptr;
  bool graph_ok = false;
  int64_t graph_steps = 0;
  void* graph_result = nullptr;
};

namespace {

size_t pbytes(int prec) { return prec == GH_F64 ? 8 : 4; }

int dalloc(void** p, size_t bytes) {
  *p = nullptr;
  if (hipMalloc(p, std::max<size_t>(bytes, 8)) != hipSuccess)
    return fail(GH_ENOMEM, "hipMalloc(%zu) failed", bytes);
  return GH_OK;
}

template <typename T>
int dupload(gh_ctx* ctx, T** d, const std::vector<T>& h) {
  GH_TRY(dalloc((void**)d, h.size() * sizeof(T)));
  if (!h.empty())
    GH_HIP(hipMemcpyAsync(*d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, ctx->stream));
  return GH_OK;
}

// ------------------------------------------------------------- V-cycle pieces
int sweep(gh_mg* mg, MgLevel& L, int niter, bool& u_zero, void* keep_u) {
  if (mg->P.smoother >= 2) {
    // StdSolver::smooth with an "ilu" smoother (the Vanka patches or the
    // ILU(0) factors): h = b - A x; solve(h); x += omega h
    const bool vanka = mg->P.smoother == 2;
    if (niter > 0 && vanka && !L.vanka) return fail(GH_EINVAL, "vanka smoother: level without patches");
    if (niter > 0 && !vanka && !L.ilu) return fail(GH_EINVAL, "ilu smoother: level without factors");
    for (int it = 0; it < niter; ++it) {
      if (!vanka) {
        // the ILU solve reads h through its order and adds omega h to x itself
        if (!u_zero) GH_TRY(csr_residual(mg->ctx, L.A, mg->wprec, L.v, L.u, L.b));
        GH_TRY(ilu_apply(mg->ctx, L.ilu, mg->wprec, u_zero ? L.b : L.v, L.u, u_zero ? 2 : 1,
                         mg->P.omega));
        u_zero = false;
        continue;
      }
      if (u_zero)
        GH_TRY(vec_convert(mg->ctx, L.size, mg->wprec, L.v, mg->wprec, L.b));
      else
        GH_TRY(csr_residual(mg->ctx, L.A, mg->wprec, L.v, L.u, L.b));
      GH_TRY(vanka_solve(mg->ctx, L.vanka, mg->wprec, L.v));
      if (u_zero)
        GH_TRY(vec_scale_to(mg->ctx, mg->wprec, L.size, L.u, mg->P.omega, L.v));
      else
        GH_TRY(vec_axpby(mg->ctx, mg->wprec, L.size, L.u, 1.0, mg->P.omega, L.v));
      u_zero = false;
    }
    (void)keep_u;
    return GH_OK;
  }
  for (int it = 0; it < niter; ++it) {
    GH_TRY(csr_jacobi_sweep(mg->ctx, L.A, mg->wprec, L.v, L.u, L.b, mg->P.omega,
                            mg->P.smoother, u_zero));
    u_zero = false;
    std::swap(L.u, L.v);
  }
  (void)keep_u;
  return GH_OK;
}

int exact(gh_mg* mg, MgLevel& L, bool u_zero) {
  if (L.direct) {
    const int64_t m = L.size;
    unsigned grid = (unsigned)((m + 3) / 4);
    if (mg->wprec == GH_F64)
      hipLaunchKernelGGL((k_dense_solve<double, double>), dim3(grid), dim3(kBlock), 0,
                         mg->ctx->stream, m, L.Ainv, (const double*)L.b, (double*)L.u);
    else
      hipLaunchKernelGGL((k_dense_solve<float, float>), dim3(grid), dim3(kBlock), 0,
                         mg->ctx->stream, m, L.Ainv, (const float*)L.b, (float*)L.u);
    GH_HIP(hipGetLastError());
    return GH_OK;
  }
  bool z = u_zero;
  GH_TRY(sweep(mg, L, mg->P.iterexact, z, nullptr));
  if (z) GH_TRY(vec_fill_zero(mg->ctx, mg->wprec, L.size, L.u));  // no sweep ran
  return GH_OK;
}

int norm_now(gh_mg* mg, const void* x, int64_t size, double* out) {
  GH_TRY(reduce_dot(mg->ctx, mg->wprec, size, x, x, 1));
  double s;
  GH_TRY(fetch_scalars(mg->ctx, 1, 1, &s));
  *out = std::sqrt(s);
  return GH_OK;
}

int correction(gh_mg* mg, int l) {
  MgLevel& F = mg->L[l];
  MgLevel& Cc = mg->L[l - 1];
  gh_ctx* ctx = mg->ctx;
  if (Cc.H) GH_TRY(hn_apply(ctx, Cc.H, Cc.ncomp, mg->wprec, Cc.u, 0));  // HNAverage(u_{l-1})
  if (F.fused_correction && F.ncomp == 1) {
    if (mg->wprec == GH_F64)
      return sell_w_launch<double>(ctx, F.p_sell, (const double*)Cc.u,
                                   CorrectOp<double, double>{(double*)F.u, mg->P.mgomega});
    return sell_w_launch<float>(ctx, F.p_sell, (const float*)Cc.u,
                                CorrectOp<float, float>{(float*)F.u, (float)mg->P.mgomega});
  }
  if (F.fused_correction) {
    unsigned grid = grid_for(F.size);
    if (mg->wprec == GH_F64)
      hipLaunchKernelGGL((k_prolongate_correct<double, double>), dim3(grid), dim3(kBlock), 0,
                         ctx->stream, F.n, F.ncomp, F.p_row, F.p_col, (const double*)Cc.u,
                         (double*)F.u, mg->P.mgomega);
    else
      hipLaunchKernelGGL((k_prolongate_correct<float, float>), dim3(grid), dim3(kBlock), 0,
                         ctx->stream, F.n, F.ncomp, F.p_row, F.p_col, (const float*)Cc.u,
                         (float*)F.u, (float)mg->P.mgomega);
    GH_HIP(hipGetLastError());
    return GH_OK;
  }
  // general path: v = 0; v += P u; HNZero(v); Dirichlet(v); u += mgomega v
  GH_TRY(vec_fill_zero(ctx, mg->wprec, F.size, F.v));
  GH_TRY(transfer_prolongate_add(ctx, F.T, F.ncomp, mg->wprec, F.v, Cc.u));
  if (F.H) GH_TRY(hn_apply(ctx, F.H, F.ncomp, mg->wprec, F.v, 2));
  if (F.M) GH_TRY(mask_apply(ctx, F.M, mg->wprec, F.v));
  return vec_axpby(ctx, mg->wprec, F.size, F.u, 1.0, mg->P.mgomega, F.v);
}

// StdMultiLevelSolver::mgstep.  need_res: compute the residual norms the
// reference computes for CGInfo (skipped when one V-cycle is all the caller
// runs: they cannot change the result).
int mgstep(gh_mg* mg, int l, int finelevel, int clevel, char& p0, char p, bool u_zero,
           bool need_res, std::vector<double>& res, std::vector<double>& rw) {
  mg->mgsteps++;
  gh_ctx* ctx = mg->ctx;
  MgLevel& L = mg->L[l];
  if (l == clevel) {
    if (p == 'F') p0 = 'V';
    GH_TRY(exact(mg, L, u_zero));
    if (l == finelevel && need_res) {
      GH_TRY(csr_residual(ctx, L.A, mg->wprec, L.v, L.u, L.b));
      GH_TRY(norm_now(mg, L.v, L.size, &res[l]));
    }
    return GH_OK;
  }
  bool z = u_zero;
  int npre = mg->P.iterpre;
  if (z && L.pre_fused) {
    // the first sweep ran in the restriction that produced b
    std::swap(L.u, L.v);
    z = false;
    npre--;
  }
  L.pre_fused = false;
  GH_TRY(sweep(mg, L, npre, z, nullptr));
  if (z) {
    // no pre-smoothing sweep ran and u is still zero: v = b
    GH_TRY(vec_convert(ctx, L.size, mg->wprec, L.v, mg->wprec, L.b));
    GH_TRY(vec_fill_zero(ctx, mg->wprec, L.size, L.u));
  } else {
    GH_TRY(csr_residual(ctx, L.A, mg->wprec, L.v, L.u, L.b));
  }
  MgLevel& Cc = mg->L[l - 1];
  // The coarse level's right-hand side is final after the restriction (no
  // hanging nodes to distribute, Dirichlet rows emptied in R) and its first
  // pre-smoothing sweep starts from zero: that sweep runs in the restriction's
  // epilogue, writing the buffer the sweep would write.
  const bool fuse = Cc.ncomp == 1 && mg->P.smoother < 2 && !Cc.H && !(Cc.M && !L.mask_in_restrict) &&
                    l - 1 != clevel && mg->P.iterpre >= 1 &&
                    (p0 == 'V' || p0 == 'W' || p0 == 'F');
  if (Cc.ncomp == 1) {
    const int rich = mg->P.smoother;
    if (fuse && mg->wprec == GH_F64)
      GH_TRY(sell_w_launch<double>(
          ctx, L.r_sell, (const double*)L.v,
          RestrictSmoothOp<double, double>{(double*)Cc.b, (double*)Cc.v,
                                           (const double*)Cc.A->dinv, mg->P.omega, rich}));
    else if (fuse)
      GH_TRY(sell_w_launch<float>(
          ctx, L.r_sell, (const float*)L.v,
          RestrictSmoothOp<float, float>{(float*)Cc.b, (float*)Cc.v, (const float*)Cc.A->dinv,
                                         (float)mg->P.omega, rich}));
    else if (mg->wprec == GH_F64)
      GH_TRY(sell_w_launch<double>(ctx, L.r_sell, (const double*)L.v,
                                   RestrictOp<double>{(double*)Cc.b}));
    else
      GH_TRY(sell_w_launch<float>(ctx, L.r_sell, (const float*)L.v,
                                  RestrictOp<float>{(float*)Cc.b}));
    Cc.pre_fused = fuse;
  } else {
    unsigned grid = grid_for(Cc.size);
    if (mg->wprec == GH_F64)
      hipLaunchKernelGGL(k_restrict<double>, dim3(grid), dim3(kBlock), 0, ctx->stream, Cc.n,
                         Cc.ncomp, L.r_row, L.r_col, (const double*)L.v, (double*)Cc.b);
    else
      hipLaunchKernelGGL(k_restrict<float>, dim3(grid), dim3(kBlock), 0, ctx->stream, Cc.n,
                         Cc.ncomp, L.r_row, L.r_col, (const float*)L.v, (float*)Cc.b);
    GH_HIP(hipGetLastError());
  }
  if (Cc.H) GH_TRY(hn_apply(ctx, Cc.H, Cc.ncomp, mg->wprec, Cc.b, 1));  // HNDistribute
  if (Cc.M && (Cc.H || !L.mask_in_restrict))
    GH_TRY(mask_apply(ctx, Cc.M, mg->wprec, Cc.b));  // SetBoundaryVectorZero
  int j = 0;
  if (p0 == 'V') j = 1;
  if (p0 == 'W') j = 2;
  if (p0 == 'F') j = 3;
  if (j == 0) GH_TRY(vec_fill_zero(ctx, mg->wprec, Cc.size, Cc.u));  // Zero(u_{l-1})
  for (int i = 0; i < j; ++i)
    GH_TRY(mgstep(mg, l - 1, finelevel, clevel, p0, p, i == 0, need_res, res, rw));
  if (l == 0 && p == 'F') p0 = 'W';
  if (need_res) GH_TRY(norm_now(mg, Cc.u, Cc.size, &rw[l]));
  GH_TRY(correction(mg, l));
  z = false;
  GH_TRY(sweep(mg, L, mg->P.iterpost, z, nullptr));
  if (need_res) {
    GH_TRY(csr_residual(ctx, L.A, mg->wprec, L.v, L.u, L.b));
    GH_TRY(norm_now(mg, L.v, L.size, &res[l]));
  }
  return GH_OK;
}

int coarsest_level(const gh_mg* mg, int finelevel, int coarselevel) {
  int clevel = coarselevel;
  for (int level = coarselevel; level < mg->nlevels; ++level)
    if (mg->L[level].direct) clevel = level;
  if (clevel > finelevel) clevel = finelevel;
  return clevel;
}

// One mgstep on the finest level (the whole of precondition() when
// prec_maxiter == 1).  The finest level reads the caller's vectors directly
// when precisions match; otherwise, and when replaying the captured graph, it
// works on staging buffers.  Level buffer roles are restored afterwards.
int fast_vcycle(gh_mg* mg, int clevel, void* xdev, int xprec, const void* ydev, int yprec,
                bool x_zero) {
  const int fl = mg->nlevels - 1;
  MgLevel& F = mg->L[fl];
  gh_ctx* ctx = mg->ctx;
  std::vector<std::pair<void*, void*>> S(mg->nlevels);
  for (int l = 0; l < mg->nlevels; ++l) S[l] = {mg->L[l].u, mg->L[l].v};
  void* save_b = F.b;
  const bool graph = mg->P.use_graph && x_zero;
  const bool direct_u = (xprec == mg->wprec) && !graph;
  const bool direct_b = (yprec == mg->wprec) && !graph;
  if (direct_b) {
    F.b = const_cast<void*>(ydev);
  } else {
    GH_TRY(vec_convert(ctx, F.size, mg->wprec, mg->fine_b, yprec, ydev));
    F.b = mg->fine_b;
  }
  // Every Jacobi / Richardson sweep of the finest level swaps its u and v
  // buffers, iterpre + iterpost times in all.  From a zero start the caller's
  // x may take the role that ends holding the result, so no copy follows.
  const bool flip = direct_u && x_zero && mg->P.smoother < 2 && fl != clevel &&
                    ((mg->P.iterpre + mg->P.iterpost) & 1);
  if (flip)
    F.v = xdev;
  else if (direct_u)
    F.u = xdev;
  else if (!x_zero)
    GH_TRY(vec_convert(ctx, F.size, mg->wprec, F.u, xprec, xdev));
  std::vector<double> res(mg->nlevels, 0.0), rw(mg->nlevels, 0.0);
  const char p = mg->P.mgtype;
  char p0 = (p == 'F') ? 'W' : p;
  void* result = nullptr;
  int rc = GH_OK;
  if (graph) {
    if (!mg->graph_ok) {
      hipGraph_t g;
      const int64_t before = mg->mgsteps;
      GH_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
      rc = mgstep(mg, fl, fl, clevel, p0, p, true, false, res, rw);
      hipError_t e = hipStreamEndCapture(ctx->stream, &g);
      if (rc) return rc;
      if (e != hipSuccess) return fail(GH_EHIP, "V-cycle capture: %s", hipGetErrorString(e));
      GH_HIP(hipGraphInstantiate(&mg->graph, g, nullptr, nullptr, 0));
      GH_HIP(hipGraphDestroy(g));
      mg->graph_steps = mg->mgsteps - before;
      mg->mgsteps = before;
      mg->graph_result = F.u;
      mg->graph_ok = true;
    }
    GH_HIP(hipGraphLaunch(mg->graph, ctx->stream));
    mg->mgsteps += mg->graph_steps;
    result = mg->graph_result;
  } else {
    rc = mgstep(mg, fl, fl, clevel, p0, p, x_zero, false, res, rw);
    result = F.u;
  }
  if (!rc && result != xdev) rc = vec_convert(ctx, F.size, xprec, xdev, mg->wprec, result);
  for (int l = 0; l < mg->nlevels; ++l) {
    mg->L[l].u = S[l].first;
    mg->L[l].v = S[l].second;
  }
  F.b = save_b;
  return rc;
}

}  // namespace

extern "C" {

}  // extern "C"

namespace {

// Renumbered copies of the caller's level operators: internal position k holds
// the caller's node perm[k].  Every row keeps its entries in the caller's order,
// so sums are formed in the same order as without the renumbering.

int build_ordered_level(gh_ctx* ctx, const gh_mg_params* params, int l, gh_csr* const* A,
                        gh_transfer* const* T, gh_hn* const* H, gh_mask* const* M,
                        const int32_t* const* perm, bool sorted64, MgLevel& Lv) {
  const gh_csr* Au = A[l];
  if (!Au->h_val_ok)
    return fail(GH_EINVAL, "gh_mg_create_ordered: level %d was assembled on the device (no host values)", l);
  const int64_t n = Au->n;
  const int B = Au->ncomp, BB = B * B;
  const int32_t* P = perm[l];
  std::vector<int32_t> Q(n, -1);
  for (int64_t k = 0; k < n; ++k) {
    if (P[k] < 0 || P[k] >= n || Q[P[k]] >= 0)
      return fail(GH_EINVAL, "gh_mg_create_ordered: order of level %d is not a permutation", l);
    Q[P[k]] = (int32_t)k;
  }
  // matrix (the adjoint solver: its transpose, formed in the caller's
  // numbering first so that rows keep the unordered solver's entry order)
  std::vector<int32_t> srp, scl;
  std::vector<double> svl;
  const int32_t* s_rp = Au->h_rowptr.data();
  const int32_t* s_col = Au->h_col.data();
  const double* s_val = Au->h_val.data();
  if (params->adjoint) {
    csr_transpose_host(n, B, s_rp, s_col, s_val, srp, scl, svl);
    s_rp = srp.data();
    s_col = scl.data();
    s_val = svl.data();
  }
  std::vector<int32_t> rp(n + 1, 0);
  for (int64_t k = 0; k < n; ++k) rp[k + 1] = rp[k] + (s_rp[P[k] + 1] - s_rp[P[k]]);
  std::vector<int32_t> cl(rp[n]);
  std::vector<double> vl((size_t)rp[n] * BB);
  parallel_rows(n, [&](int64_t a, int64_t b) {
    for (int64_t k = a; k < b; ++k) {
      const int32_t o = P[k];
      int64_t t = rp[k];
      for (int32_t q = s_rp[o]; q < s_rp[o + 1]; ++q, ++t) {
        cl[t] = Q[s_col[q]];
        std::copy(s_val + (int64_t)q * BB, s_val + (int64_t)(q + 1) * BB, vl.begin() + t * BB);
      }
    }
  });
  GH_TRY(csr_create_impl(ctx, n, B, rp.data(), cl.data(), vl.data(), GH_F64, false, &Lv.A64));
  Lv.own_A64 = true;
  if (params->prec != GH_F64 || sorted64) {
    // The reduced-precision V-cycle reproduces no reference bits, so its rows
    // take their entries in ascending internal column order: rows of one kind
    // then share one relative pattern, whatever order the caller's numbering
    // gave them.  The same holds for the outer FP64 products of flexible GMRES.
    std::vector<int32_t> cs(cl.size());
    std::vector<double> vs(vl.size());
    parallel_rows(n, [&](int64_t a, int64_t b) {
      std::vector<int32_t> idx;
      for (int64_t k = a; k < b; ++k) {
        idx.resize(rp[k + 1] - rp[k]);
        for (size_t t = 0; t < idx.size(); ++t) idx[t] = rp[k] + (int32_t)t;
        std::sort(idx.begin(), idx.end(), [&](int32_t x, int32_t y) { return cl[x] < cl[y]; });
        for (size_t t = 0; t < idx.size(); ++t) {
          cs[rp[k] + t] = cl[idx[t]];
          std::copy(vl.begin() + (int64_t)idx[t] * BB, vl.begin() + (int64_t)(idx[t] + 1) * BB,
                    vs.begin() + (int64_t)(rp[k] + t) * BB);
        }
      }
    });
    if (sorted64)
      GH_TRY(csr_create_impl(ctx, n, B, rp.data(), cs.data(), vs.data(), GH_F64, false, &Lv.A64s));
    if (params->prec != GH_F64) {
      GH_TRY(csr_create_impl(ctx, n, B, rp.data(), cs.data(), vs.data(), params->prec, false, &Lv.A));
      Lv.own_A = true;
    } else {
      Lv.A = Lv.A64;
    }
  } else {
    Lv.A = Lv.A64;
  }
  std::vector<int32_t> hperm(P, P + n);
  GH_TRY(dupload(ctx, &Lv.d_perm, hperm));
  // transfer to level l-1
  if (l > 0) {
    const gh_transfer* Tu = T[l];
    const int64_t nc = Tu->ncoarse;
    const int32_t* Pc = perm[l - 1];
    std::vector<int32_t> Qc(nc);
    for (int64_t k = 0; k < nc; ++k) Qc[Pc[k]] = (int32_t)k;
    std::vector<int32_t> rrow(nc + 1, 0), prow(n + 1, 0);
    std::vector<uint32_t> rcol, pcol;
    rcol.reserve(Tu->h_rcol.size());
    pcol.reserve(Tu->h_pcol.size());
    for (int64_t k = 0; k < nc; ++k) {
      const int32_t o = Pc[k];
      for (int32_t q = Tu->h_rrow[o]; q < Tu->h_rrow[o + 1]; ++q) {
        const uint32_t e = Tu->h_rcol[q];
        rcol.push_back((uint32_t)Q[e & kIdxMask] | (e & ~kIdxMask));
      }
      rrow[k + 1] = (int32_t)rcol.size();
    }
    for (int64_t k = 0; k < n; ++k) {
      const int32_t o = P[k];
      for (int32_t q = Tu->h_prow[o]; q < Tu->h_prow[o + 1]; ++q) {
        const uint32_t e = Tu->h_pcol[q];
        pcol.push_back((uint32_t)Qc[e & kIdxMask] | (e & ~kIdxMask));
      }
      prow[k + 1] = (int32_t)pcol.size();
    }
    GH_TRY(transfer_from_csr(ctx, nc, n, std::move(rrow), std::move(rcol), std::move(prow),
                             std::move(pcol), &Lv.T));
    Lv.own_T = true;
  }
  auto mapv = [&](int32_t v) { return (v >= 0 && v < n) ? Q[v] : v; };
  if (H && H[l]) {
    std::vector<int32_t> e(H[l]->h_edges), f(H[l]->h_faces);
    for (int32_t& v : e) v = mapv(v);
    for (int32_t& v : f) v = mapv(v);
    GH_TRY(gh_hn_create(ctx, H[l]->kind, n, H[l]->ne, e.data(), H[l]->nf, f.data(), &Lv.H));
    Lv.own_H = true;
  }
  if (M && M[l]) {
    std::vector<int32_t> nodes(M[l]->h_nodes);
    for (int32_t& v : nodes) v = Q[v];
    GH_TRY(gh_mask_create(ctx, n, B, (int64_t)nodes.size(), nodes.data(), M[l]->ncomps,
                          M[l]->h_comps.data(), &Lv.M));
    Lv.own_M = true;
  }
  return GH_OK;
}

int mg_build(gh_ctx* ctx, int nlevels, gh_csr* const* A, gh_transfer* const* T, gh_hn* const* H,
             gh_mask* const* M, const int32_t* const* perm, const gh_mg_params* params,
             gh_mg** out) {
  GH_ARG(ctx && out && nlevels >= 1 && A && params, "gh_mg_create: bad arguments");
  GH_ARG(params->prec == GH_F64 || params->prec == GH_F32 || params->prec == GH_F16,
         "gh_mg_create: bad precision");
  GH_ARG(params->mgtype == 'V' || params->mgtype == 'W' || params->mgtype == 'F',
         "gh_mg_create: mgtype must be V, W or F");
  GH_ARG(params->smoother >= 0 && params->smoother <= 3, "gh_mg_create: bad smoother");
  GH_ARG(params->krylov == 0 || params->krylov == 1, "gh_mg_create: krylov must be 0 or 1");
  GH_ARG(params->adjoint == 0 || params->adjoint == 1, "gh_mg_create: adjoint must be 0 or 1");
  // shapes of the caller's objects
  for (int l = 0; l < nlevels; ++l) {
    if (!A[l] || A[l]->prec != GH_F64)
      return fail(GH_EINVAL, "gh_mg_create: level %d needs an FP64 matrix", l);
    if (!A[l]->h_val_ok && (params->adjoint || params->prec != GH_F64 || perm))
      return fail(GH_EINVAL,
                  "gh_mg_create: level %d was assembled on the device; its copies cannot be formed",
                  l);
    if (H && H[l] && H[l]->n != A[l]->n)
      return fail(GH_EINVAL, "gh_mg_create: HN size mismatch on level %d", l);
    if (M && M[l] && (M[l]->n != A[l]->n || M[l]->ncomp != A[l]->ncomp))
      return fail(GH_EINVAL, "gh_mg_create: mask size mismatch on level %d", l);
    if (l > 0) {
      if (!T || !T[l]) return fail(GH_EINVAL, "gh_mg_create: level %d has no transfer", l);
      if (T[l]->nfine != A[l]->n || T[l]->ncoarse != A[l - 1]->n || A[l - 1]->ncomp != A[l]->ncomp)
        return fail(GH_EINVAL, "gh_mg_create: transfer %d does not match the level sizes", l);
    }
    if (perm && !perm[l]) return fail(GH_EINVAL, "gh_mg_create_ordered: level %d has no order", l);
  }
  *out = nullptr;
  gh_mg* mg = new gh_mg();
  mg->ctx = ctx;
  mg->nlevels = nlevels;
  mg->P = *params;
  mg->wprec = params->prec == GH_F64 ? GH_F64 : GH_F32;
  mg->L.resize(nlevels);
  mg->ordered = perm != nullptr;
  mg->user_fine = A[nlevels - 1];
  mg->user_mask = M ? M[nlevels - 1] : nullptr;
  mg->user_hn = H ? H[nlevels - 1] : nullptr;
  ctx_retain(ctx);  // balanced by gh_mg_destroy
  int rc = GH_OK;
  for (int l = 0; l < nlevels && !rc; ++l) {
    MgLevel& Lv = mg->L[l];
    Lv.n = A[l]->n;
    Lv.ncomp = A[l]->ncomp;
    Lv.size = Lv.n * Lv.ncomp;
    if (perm) {
      rc = build_ordered_level(ctx, params, l, A, T, H, M, perm,
                               params->krylov == 1 && l == nlevels - 1, Lv);
    } else if (params->adjoint) {
      // the transposed level operator, once; everything downstream runs on it
      std::vector<int32_t> rpT, clT;
      std::vector<double> vlT;
      csr_transpose_host(A[l]->n, A[l]->ncomp, A[l]->h_rowptr.data(), A[l]->h_col.data(),
                         A[l]->h_val.data(), rpT, clT, vlT);
      rc = csr_create_impl(ctx, Lv.n, Lv.ncomp, rpT.data(), clT.data(), vlT.data(), GH_F64, false,
                           &Lv.A64);
      Lv.own_A64 = !rc;
      if (!rc && params->prec == GH_F64) {
        Lv.A = Lv.A64;
      } else if (!rc) {
        rc = csr_create_impl(ctx, Lv.n, Lv.ncomp, rpT.data(), clT.data(), vlT.data(), params->prec,
                             false, &Lv.A);
        Lv.own_A = !rc;
      }
      Lv.T = l > 0 ? T[l] : nullptr;
      // a condensed level operator has identity-only hanging rows and columns,
      // so its transpose is condensed over the same constraints: the V-cycle
      // keeps the forward HNAverage / HNDistribute / HNZero
      Lv.H = H ? H[l] : nullptr;
      Lv.M = M ? M[l] : nullptr;
    } else {
      Lv.A64 = A[l];
      Lv.T = l > 0 ? T[l] : nullptr;
      Lv.H = H ? H[l] : nullptr;
      Lv.M = M ? M[l] : nullptr;
      if (params->prec == GH_F64) {
        Lv.A = A[l];
      } else {
        rc = gh_csr_create(ctx, A[l]->n, A[l]->ncomp, A[l]->h_rowptr.data(), A[l]->h_col.data(),
                           A[l]->h_val.data(), params->prec, &Lv.A);
        Lv.own_A = true;
      }
    }
    if (!rc) rc = dalloc(&Lv.u, Lv.size * pbytes(mg->wprec));
    if (!rc) rc = dalloc(&Lv.b, Lv.size * pbytes(mg->wprec));
    if (!rc) rc = dalloc(&Lv.v, Lv.size * pbytes(mg->wprec));
    if (!rc) rc = vec_fill_zero(ctx, mg->wprec, Lv.size, Lv.u);
    if (!rc) rc = vec_fill_zero(ctx, mg->wprec, Lv.size, Lv.b);
    if (!rc) rc = vec_fill_zero(ctx, mg->wprec, Lv.size, Lv.v);
    Lv.direct = Lv.n < params->ndirect;
    if (!rc && Lv.direct) {
      rc = dalloc((void**)&Lv.Ainv, sizeof(double) * Lv.size * Lv.size);
      if (!rc) rc = dense_inverse_device(ctx, Lv.A64, Lv.Ainv);
    }
  }
  // transfer operators with the constraints folded in
  for (int l = 1; l < nlevels && !rc; ++l) {
    MgLevel& F = mg->L[l];
    MgLevel& C = mg->L[l - 1];
    const gh_transfer* Tr = F.T;
    const int B = F.ncomp;
    // restriction rows of coarse Dirichlet nodes (all components) emptied
    std::vector<uint8_t> cdir(C.n, 0);
    F.mask_in_restrict = (C.M == nullptr) || (C.M->ncomps == B);
    if (C.M && F.mask_in_restrict)
      for (int32_t v : C.M->h_nodes) cdir[v] = 1;
    std::vector<int32_t> rrow(1, 0);
    std::vector<uint32_t> rcol;
    for (int64_t i = 0; i < C.n; ++i) {
      if (!cdir[i])
        for (int32_t q = Tr->h_rrow[i]; q < Tr->h_rrow[i + 1]; ++q) rcol.push_back(Tr->h_rcol[q]);
      rrow.push_back((int32_t)rcol.size());
    }
    // prolongation rows of fine hanging / Dirichlet nodes emptied
    F.fused_correction = (F.M == nullptr) || (F.M->ncomps == B);
    std::vector<uint8_t> fz(F.n, 0);
    if (F.M && F.fused_correction)
      for (int32_t v : F.M->h_nodes) fz[v] = 1;
    if (F.H) {
      for (int64_t k = 0; k < F.H->ne; ++k) fz[F.H->h_edges[4 * k]] = 1;
      for (int64_t k = 0; k < F.H->nf; ++k) fz[F.H->h_faces[10 * k]] = 1;
    }
    std::vector<int32_t> prow(1, 0);
    std::vector<uint32_t> pcol;
    for (int64_t i = 0; i < F.n; ++i) {
      if (!fz[i])
        for (int32_t q = Tr->h_prow[i]; q < Tr->h_prow[i + 1]; ++q) pcol.push_back(Tr->h_pcol[q]);
      prow.push_back((int32_t)pcol.size());
    }
    rc = dupload(ctx, &F.r_row, rrow);
    if (!rc) rc = dupload(ctx, &F.r_col, rcol);
    if (!rc) rc = dupload(ctx, &F.p_row, prow);
    if (!rc) rc = dupload(ctx, &F.p_col, pcol);
    if (!rc) rc = sell_build(ctx, C.n, rrow.data(), rcol.data(), nullptr, GH_F64, &F.r_sell);
    if (!rc) rc = sell_build(ctx, F.n, prow.data(), pcol.data(), nullptr, GH_F64, &F.p_sell);
  }
  const int64_t nf = mg->L[nlevels - 1].size;
  if (!rc) rc = dalloc(&mg->fine_b, nf * pbytes(mg->wprec));
  if (!rc && mg->ordered) rc = dalloc((void**)&mg->ext_x, nf * 8);
  if (!rc && mg->ordered) rc = dalloc((void**)&mg->ext_y, nf * 8);
  if (!rc && hipStreamSynchronize(ctx->stream) != hipSuccess) rc = fail(GH_EHIP, "sync");
  if (rc) {
    gh_mg_destroy(mg);
    return rc;
  }
  *out = mg;
  return GH_OK;
}

}  // namespace

extern "C" {

int gh_mg_create(gh_ctx* ctx, int nlevels, gh_csr* const* A, gh_transfer* const* T,
                 gh_hn* const* H, gh_mask* const* M, const gh_mg_params* params,
                 gh_mg** out) {
  GH_ARG(A && nlevels >= 1, "gh_mg_create: bad arguments");
  for (int l = 0; l < nlevels; ++l)
    GH_ARG(!A[l] || !A[l]->host_only,
           "gh_mg_create: level matrices need a device layout (host-only matrices serve "
           "gh_mg_create_ordered)");
  return mg_build(ctx, nlevels, A, T, H, M, nullptr, params, out);
}

int gh_mg_create_ordered(gh_ctx* ctx, int nlevels, gh_csr* const* A, gh_transfer* const* T,
                         gh_hn* const* H, gh_mask* const* M, const int32_t* const* order,
                         const gh_mg_params* params, gh_mg** out) {
  GH_ARG(order, "gh_mg_create_ordered: order is NULL");
  return mg_build(ctx, nlevels, A, T, H, M, order, params, out);
}

int gh_mg_destroy(gh_mg* mg) {
  if (!mg) return GH_OK;
  if (mg->ctx && mg->ctx->stream) (void)hipStreamSynchronize(mg->ctx->stream);
  for (MgLevel& L : mg->L) {
    if (L.own_A && L.A) gh_csr_destroy(L.A);
    if (L.own_A64 && L.A64) gh_csr_destroy(L.A64);
    if (L.A64s) gh_csr_destroy(L.A64s);
    if (L.own_T && L.T) gh_transfer_destroy(L.T);
    if (L.own_H && L.H) gh_hn_destroy(L.H);
    if (L.own_M && L.M) gh_mask_destroy(L.M);
    if (L.d_perm) (void)hipFree(L.d_perm);
    if (L.vanka) gh_vanka_destroy(L.vanka);
    if (L.ilu) gh_ilu_destroy(L.ilu);
    for (void* p : {(void*)L.r_row, (void*)L.r_col, (void*)L.p_row, (void*)L.p_col, L.u, L.b,
                    L.v, (void*)L.Ainv})
      if (p) (void)hipFree(p);
    sell_free(&L.r_sell);
    sell_free(&L.p_sell);
  }
  for (double* m : mg->mem)
    if (m) (void)hipFree(m);
  for (void* z : mg->zmem)
    if (z) (void)hipFree(z);
  if (mg->d_basis) (void)hipFree((void*)mg->d_basis);
  if (mg->d_coef) (void)hipFree(mg->d_coef);
  if (mg->h_coef) (void)hipHostFree(mg->h_coef);
  for (void* p : {(void*)mg->p, (void*)mg->r, (void*)mg->w, mg->fine_b, (void*)mg->ext_x,
                  (void*)mg->ext_y})
    if (p) (void)hipFree(p);
  if (mg->graph) (void)hipGraphExecDestroy(mg->graph);
  gh_ctx* ctx = mg->ctx;
  delete mg;
  ctx_release(ctx);
  return GH_OK;
}

int gh_mg_assemble_q1(gh_ctx* ctx, gh_mg* mg, gh_cells* const* C, gh_vec* const* k) {
  GH_ARG(ctx && mg && C, "gh_mg_assemble_q1: NULL argument");
  GH_ARG(!mg->P.adjoint, "gh_mg_assemble_q1: the adjoint solver's transposed levels are not assembled");
  GH_ARG(mg->P.smoother < 2, "gh_mg_assemble_q1: Vanka and ILU factors are not re-formed");
  for (int l = 0; l < mg->nlevels; ++l) {
    const MgLevel& L = mg->L[l];
    if (!C[l] || C[l]->n != L.n || L.ncomp != 1)
      return fail(GH_EINVAL, "gh_mg_assemble_q1: cells of level %d do not match a scalar level", l);
    if (k && k[l] && (k[l]->prec != GH_F64 || k[l]->size != C[l]->ncells))
      return fail(GH_EINVAL, "gh_mg_assemble_q1: k of level %d needs one FP64 entry per cell", l);
    if (L.H)
      return fail(GH_EINVAL, "gh_mg_assemble_q1: level %d has hanging nodes (not assembled here)", l);
  }
  for (int l = 0; l < mg->nlevels; ++l) {
    MgLevel& L = mg->L[l];
    const uint8_t* isdir = nullptr;
    if (L.M) GH_TRY(mask_node_flags(ctx, L.M, &isdir));
    const double* kc = (k && k[l]) ? (const double*)k[l]->d : nullptr;
    gh_csr* mats[3] = {L.A64, L.A != L.A64 ? L.A : nullptr, L.A64s};
    for (gh_csr* M : mats) {
      if (!M) continue;
      double* v64 = nullptr;
      GH_TRY(csr_asm_values(ctx, M, &v64));
      GH_TRY(assemble_q1_launch(ctx, C[l], kc, isdir, L.d_perm, M, v64));
      GH_TRY(csr_asm_commit(ctx, M, "gh_mg_assemble_q1"));
    }
    if (L.direct) GH_TRY(dense_inverse_device(ctx, L.A64, L.Ainv));
  }
  GH_HIP(hipStreamSynchronize(ctx->stream));
  return GH_OK;
}

int gh_mg_assemble_q2(gh_ctx* ctx, gh_mg* mg, gh_patches* const* P, gh_vec* const* k) {
  GH_ARG(ctx && mg && P, "gh_mg_assemble_q2: NULL argument");
  GH_ARG(!mg->P.adjoint, "gh_mg_assemble_q2: the adjoint solver's transposed levels are not assembled");
  GH_ARG(mg->P.smoother < 2, "gh_mg_assemble_q2: Vanka and ILU factors are not re-formed");
  for (int l = 0; l < mg->nlevels; ++l) {
    const MgLevel& L = mg->L[l];
    if (!P[l] || P[l]->dim != 3 || P[l]->nnodes != L.n || L.ncomp != 1)
      return fail(GH_EINVAL, "gh_mg_assemble_q2: patches of level %d do not match a scalar level", l);
    if (k && k[l] && (k[l]->prec != GH_F64 || k[l]->size != P[l]->npatches))
      return fail(GH_EINVAL, "gh_mg_assemble_q2: k of level %d needs one FP64 entry per patch", l);
    if (L.H)
      return fail(GH_EINVAL, "gh_mg_assemble_q2: level %d has hanging nodes (not assembled here)", l);
  }
  for (int l = 0; l < mg->nlevels; ++l) {
    MgLevel& L = mg->L[l];
    const uint8_t* isdir = nullptr;
    if (L.M) GH_TRY(mask_node_flags(ctx, L.M, &isdir));
    // renumbered level: the order and its inverse (node -> position)
    std::vector<int32_t> perm_h, iperm_h;
    int32_t* d_iperm = nullptr;
    if (L.d_perm) {
      perm_h.resize(L.n);
      iperm_h.resize(L.n);
      GH_HIP(hipMemcpy(perm_h.data(), L.d_perm, sizeof(int32_t) * L.n, hipMemcpyDeviceToHost));
      for (int64_t i = 0; i < L.n; ++i) iperm_h[perm_h[i]] = (int32_t)i;
      GH_TRY(dupload(ctx, &d_iperm, iperm_h));
    }
    const double* kp = (k && k[l]) ? (const double*)k[l]->d : nullptr;
    gh_csr* mats[3] = {L.A64, L.A != L.A64 ? L.A : nullptr, L.A64s};
    int rc = GH_OK;
    for (gh_csr* M : mats) {
      if (!M || rc) continue;
      double* v64 = nullptr;
      rc = csr_asm_values(ctx, M, &v64);
      const int mode = rc ? 0 : assemble_q2_mode(M, L.d_perm ? &perm_h : nullptr);
      if (!rc) rc = assemble_q2_launch(ctx, P[l], kp, isdir, L.d_perm, d_iperm, mode, M, v64);
      if (!rc) rc = csr_asm_commit(ctx, M, "gh_mg_assemble_q2");
    }
    if (d_iperm) {
      (void)hipStreamSynchronize(ctx->stream);
      (void)hipFree(d_iperm);
    }
    GH_TRY(rc);
    if (L.direct) GH_TRY(dense_inverse_device(ctx, L.A64, L.Ainv));
  }
  GH_HIP(hipStreamSynchronize(ctx->stream));
  return GH_OK;
}

int gh_mg_set_vanka(gh_mg* mg, int level, int64_t npatches, int psize, const int32_t* nodes) {
  GH_ARG(mg && level >= 0 && level < mg->nlevels && (nodes || npatches == 0) && npatches >= 0 &&
             psize >= 1,
         "gh_mg_set_vanka: bad arguments");
  MgLevel& L = mg->L[level];
  std::vector<int32_t> internal(nodes, nodes + npatches * psize);
  if (L.d_perm) {
    // caller's node -> internal position
    std::vector<int32_t> P(L.n), Q(L.n);
    GH_HIP(hipMemcpy(P.data(), L.d_perm, sizeof(int32_t) * L.n, hipMemcpyDeviceToHost));
    for (int64_t k = 0; k < L.n; ++k) Q[P[k]] = (int32_t)k;
    for (int32_t& v : internal) {
      if (v < 0 || v >= L.n) return fail(GH_EINVAL, "gh_mg_set_vanka: node out of range");
      v = Q[v];
    }
  }
  gh_vanka* V = nullptr;
  GH_TRY(vanka_create(mg->ctx, L.A64, npatches, psize, internal.data(), mg->wprec, &V));
  if (L.vanka) gh_vanka_destroy(L.vanka);
  L.vanka = V;
  if (mg->graph) (void)hipGraphExecDestroy(mg->graph);  // it holds the old smoother
  mg->graph = nullptr;
  mg->graph_ok = false;
  return GH_OK;
}

int gh_mg_set_ilu(gh_mg* mg, int level, const int32_t* perm) {
  GH_ARG(mg && level >= 0 && level < mg->nlevels, "gh_mg_set_ilu: bad arguments");
  MgLevel& L = mg->L[level];
  // ILU row i is the caller's node perm[i]: in the level's internal numbering
  // the node at internal position Q[perm[i]]
  std::vector<int32_t> internal(L.n);
  std::vector<int32_t> P(L.n), Q(L.n);
  if (L.d_perm) {
    GH_HIP(hipMemcpy(P.data(), L.d_perm, sizeof(int32_t) * L.n, hipMemcpyDeviceToHost));
    for (int64_t k = 0; k < L.n; ++k) Q[P[k]] = (int32_t)k;
  } else {
    for (int64_t k = 0; k < L.n; ++k) Q[k] = (int32_t)k;
  }
  for (int64_t i = 0; i < L.n; ++i) {
    const int32_t v = perm ? perm[i] : (int32_t)i;
    if (v < 0 || v >= L.n) return fail(GH_EINVAL, "gh_mg_set_ilu: node out of range");
    internal[i] = Q[v];
  }
  gh_ilu* I = nullptr;
  GH_TRY(gh_ilu_create(mg->ctx, L.A64, internal.data(), &I));
  if (L.ilu) gh_ilu_destroy(L.ilu);
  L.ilu = I;
  if (mg->graph) (void)hipGraphExecDestroy(mg->graph);  // it holds the old smoother
  mg->graph = nullptr;
  mg->graph_ok = false;
  return GH_OK;
}

int gh_mg_nlevels(const gh_mg* mg) { return mg ? mg->nlevels : -1; }
int64_t gh_mg_mgsteps(const gh_mg* mg) { return mg ? mg->mgsteps : -1; }
const gh_csr* gh_mg_level_matrix(const gh_mg* mg, int level, int fp64) {
  if (!mg || level < 0 || level >= mg->nlevels) return nullptr;
  return fp64 ? mg->L[level].A64 : mg->L[level].A;
}

int gh_mg_linear(gh_ctx* ctx, gh_mg* mg, int finelevel, int coarselevel, gh_vec* u,
                 const gh_vec* f, gh_cginfo* info) {
  GH_ARG(ctx && mg && u && f && info, "gh_mg_linear: NULL argument");
  GH_ARG(finelevel >= 0 && finelevel < mg->nlevels && coarselevel >= 0,
         "gh_mg_linear: bad levels");
  MgLevel& F = mg->L[finelevel];
  GH_ARG(u->n == F.n && f->n == F.n && u->ncomp == F.ncomp && f->ncomp == F.ncomp,
         "gh_mg_linear: vector shape does not match the level");
  GH_ARG(u->d != f->d, "gh_mg_linear: u and f must not alias");
  const int clevel = coarsest_level(mg, finelevel, coarselevel);
  // level vectors: u, b = _mg0_ (= f), v = _mg1_ in working precision
  if (F.d_perm) {
    GH_TRY(vec_permute(ctx, F.n, F.ncomp, F.d_perm, mg->wprec, F.u, u->prec, u->d, true));
    GH_TRY(vec_permute(ctx, F.n, F.ncomp, F.d_perm, mg->wprec, F.b, f->prec, f->d, true));
  } else {
    GH_TRY(vec_convert(ctx, F.size, mg->wprec, F.u, u->prec, u->d));
    GH_TRY(vec_convert(ctx, F.size, mg->wprec, F.b, f->prec, f->d));
  }
  std::vector<double> res(mg->nlevels, 0.0), rw(mg->nlevels, 0.0);
  GH_TRY(csr_residual(ctx, F.A, mg->wprec, F.v, F.u, F.b));
  GH_TRY(norm_now(mg, F.v, F.size, &res[finelevel]));
  rw[finelevel] = 0.;
  gh_cginfo_check(info, res[finelevel], rw[finelevel]);
  int reached = 0;
  while (!reached) {
    char p = mg->P.mgtype;
    char p0 = (p == 'F') ? 'W' : p;
    GH_TRY(mgstep(mg, finelevel, finelevel, clevel, p0, p, false, true, res, rw));
    reached = gh_cginfo_check(info, res[finelevel], rw[finelevel]);
  }
  if (F.d_perm)
    GH_TRY(vec_permute(ctx, F.n, F.ncomp, F.d_perm, u->prec, u->d, mg->wprec, F.u, false));
  else
    GH_TRY(vec_convert(ctx, F.size, u->prec, u->d, mg->wprec, F.u));
  GH_HIP(hipStreamSynchronize(ctx->stream));
  // the level's buffers may have swapped roles; all three are scratch
  return GH_OK;
}

int gh_mg_precondition(gh_ctx* ctx, gh_mg* mg, int coarselevel, gh_vec* x, const gh_vec* y,
                       int x_is_zero) {
  GH_ARG(ctx && mg && x && y, "gh_mg_precondition: NULL argument");
  const int fl = mg->nlevels - 1;
  MgLevel& F = mg->L[fl];
  GH_ARG(x->n == F.n && y->n == F.n && x->ncomp == F.ncomp && y->ncomp == F.ncomp,
         "gh_mg_precondition: vector shape does not match the finest level");
  GH_ARG(x->d != y->d, "gh_mg_precondition: x and y must not alias");
  const int cl = coarselevel < 0 ? fl : coarselevel;
  if (mg->P.prec_maxiter != 1) {
    // CGInfo with prec_maxiter > 1 needs the residual norms: general path
    gh_cginfo prec;
    gh_cginfo_init(&prec, mg->P.prec_tol, mg->P.prec_globaltol, mg->P.prec_maxiter);
    gh_cginfo_reset(&prec);
    gh_cginfo_check(&prec, 0., 0.);
    return gh_mg_linear(ctx, mg, fl, cl, x, y, &prec);
  }
  // prec_maxiter == 1: precondition() always runs exactly one mgstep
  // (CGInfo::check re-initialises and stops with "too many iterations")
  const int clevel = coarsest_level(mg, fl, cl);
  if (!F.d_perm) return fast_vcycle(mg, clevel, x->d, x->prec, y->d, y->prec, x_is_zero != 0);
  GH_TRY(vec_permute(ctx, F.n, F.ncomp, F.d_perm, GH_F64, mg->ext_y, y->prec, y->d, true));
  if (!x_is_zero)
    GH_TRY(vec_permute(ctx, F.n, F.ncomp, F.d_perm, GH_F64, mg->ext_x, x->prec, x->d, true));
  GH_TRY(fast_vcycle(mg, clevel, mg->ext_x, GH_F64, mg->ext_y, GH_F64, x_is_zero != 0));
  return vec_permute(ctx, F.n, F.ncomp, F.d_perm, x->prec, x->d, GH_F64, mg->ext_x, false);
}

}  // extern "C"

// ==================================================================== GMRES
namespace {

// x += sum_k h_k vs[k], the terms added in order k = 0..m-1 for every entry
// (GMRES::solution, gmres.cc:235-260, is one Add per basis vector).
template <typename Z>
__global__ __launch_bounds__(kBlock) void k_multi_axpy(int64_t n, double* __restrict__ x, int m,
                                                       const Z* const* __restrict__ vs,
                                                       const double* __restrict__ h) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock) {
    double acc = x[i];
    for (int k = 0; k < m; ++k) acc = add_rn(acc, mul_rn(h[k], (double)vs[k][i]));
    x[i] = acc;
  }
}

constexpr int kMaxBasis = 256;  // vmax <= 256

// GMRES memory of the finest level: mem[0..vmax] (FP64 Arnoldi basis), with
// `flexible` also zmem[0..vmax] (preconditioned directions in the V-cycle's
// precision), and the pointer table the solution update reads.
int ensure_gmres_memory(gh_mg* mg, int vmax, bool flexible) {
  const int64_t n = mg->L[mg->nlevels - 1].size;
  bool grown = false;
  if ((int)mg->mem.size() < vmax + 1) mg->mem.resize(vmax + 1, nullptr);
  for (int i = 0; i <= vmax; ++i)
    if (!mg->mem[i]) {
      GH_TRY(dalloc((void**)&mg->mem[i], n * 8));
      grown = true;
    }
  if (flexible) {
    if ((int)mg->zmem.size() < vmax + 1) mg->zmem.resize(vmax + 1, nullptr);
    for (int i = 0; i <= vmax; ++i)
      if (!mg->zmem[i]) {
        GH_TRY(dalloc(&mg->zmem[i], n * pbytes(mg->wprec)));
        grown = true;
      }
  }
  if (!mg->p) GH_TRY(dalloc((void**)&mg->p, n * 8));
  if (!mg->r) GH_TRY(dalloc((void**)&mg->r, n * 8));
  if (!mg->w) GH_TRY(dalloc((void**)&mg->w, n * 8));
  if (!mg->d_coef) GH_TRY(dalloc((void**)&mg->d_coef, kMaxBasis * 8));
  if (!mg->h_coef) GH_HIP(hipHostMalloc((void**)&mg->h_coef, kMaxBasis * 8, hipHostMallocDefault));
  const int cap = (int)(mg->mem.size() + mg->zmem.size());
  if (grown || !mg->d_basis || mg->basis_cap != cap) {
    if (mg->d_basis) (void)hipFree((void*)mg->d_basis);
    mg->d_basis = nullptr;
    GH_TRY(dalloc((void**)&mg->d_basis, sizeof(void*) * cap));
    std::vector<const void*> t;
    for (double* m : mg->mem) t.push_back(m);
    for (void* z : mg->zmem) t.push_back(z);
    GH_HIP(hipMemcpy((void*)mg->d_basis, t.data(), sizeof(void*) * cap, hipMemcpyHostToDevice));
    mg->basis_cap = cap;
  }
  return GH_OK;
}

// x += sum_{i<dim} c_i basis_i with basis mem (FP64) or zmem (V-cycle precision).
// The coefficients go through pinned memory; every caller synchronises the
// stream (a fetched scalar) before the next update rewrites them.
int basis_update(gh_mg* mg, double* x, bool zbasis, int dim, const double* c) {
  gh_ctx* ctx = mg->ctx;
  const int64_t n = mg->L[mg->nlevels - 1].size;
  for (int i = 0; i < dim; ++i) mg->h_coef[i] = c[i];
  GH_HIP(hipMemcpyAsync(mg->d_coef, mg->h_coef, sizeof(double) * dim, hipMemcpyHostToDevice,
                        ctx->stream));
  const dim3 grid((unsigned)std::min<int64_t>((n + kBlock - 1) / kBlock, 8192));
  const void* const* tab = mg->d_basis + (zbasis ? mg->mem.size() : 0);
  if (!zbasis || mg->wprec == GH_F64)
    hipLaunchKernelGGL(k_multi_axpy<double>, grid, dim3(kBlock), 0, ctx->stream, n, x, dim,
                       (const double* const*)tab, (const double*)mg->d_coef);
  else
    hipLaunchKernelGGL(k_multi_axpy<float>, grid, dim3(kBlock), 0, ctx->stream, n, x, dim,
                       (const float* const*)tab, (const double*)mg->d_coef);
  GH_HIP(hipGetLastError());
  return GH_OK;
}

int precond_raw(gh_mg* mg, int coarselevel, double* x, const double* y) {
  const int fl = mg->nlevels - 1;
  const int cl = coarselevel < 0 ? fl : coarselevel;
  if (mg->P.prec_maxiter != 1) {
    // internal numbering: run LinearMg without the boundary permutation
    gh_vec xv, yv;
    xv.ctx = yv.ctx = mg->ctx;
    xv.n = yv.n = mg->L[fl].n;
    xv.ncomp = yv.ncomp = mg->L[fl].ncomp;
    xv.size = yv.size = mg->L[fl].size;
    xv.prec = yv.prec = GH_F64;
    xv.d = x;
    yv.d = const_cast<double*>(y);
    GH_TRY(vec_fill_zero(mg->ctx, GH_F64, xv.size, x));
    gh_cginfo prec;
    gh_cginfo_init(&prec, mg->P.prec_tol, mg->P.prec_globaltol, mg->P.prec_maxiter);
    gh_cginfo_reset(&prec);
    gh_cginfo_check(&prec, 0., 0.);
    int32_t* saved = mg->L[fl].d_perm;
    mg->L[fl].d_perm = nullptr;
    int rc = gh_mg_linear(mg->ctx, mg, fl, cl, &xv, &yv, &prec);
    mg->L[fl].d_perm = saved;
    return rc;
  }
  const int clevel = coarsest_level(mg, fl, cl);
  return fast_vcycle(mg, clevel, x, GH_F64, y, GH_F64, true);
}

// Arnoldi step shared by both GMRES forms (gmres.cc:109-137, 91-105): modified
// Gram-Schmidt of vv against mem[0..dim-1] into h (dots in slots 8..), the
// re-orthogonalisation test, normalisation, and the Givens rotation of column
// `dim-1`.  Returns |gamma[dim]|.
int arnoldi_step(gh_mg* mg, int dim, double* vv, std::vector<double>& h,
                 std::vector<double>& gamma, std::vector<double>& ci, std::vector<double>& si,
                 std::vector<double>& d, double* rho, bool classical = false) {
  gh_ctx* ctx = mg->ctx;
  const int64_t n = mg->L[mg->nlevels - 1].size;
  const int last = dim - 1;
  if (classical) {
    // flexible GMRES (no reference sums to keep): classical Gram-Schmidt,
    // h_k = <vv, v_k> for all k at once, then vv -= sum_k h_k v_k
    const double* const* basis = (const double* const*)mg->d_basis;
    GH_TRY(cgs_dots(ctx, n, vv, basis, dim, 8));
    GH_TRY(cgs_update(ctx, n, vv, basis, dim, 8, 8 + dim));
    GH_TRY(fetch_scalars(ctx, 8, dim + 1, d.data()));
    for (int i = 0; i < dim; ++i) h[i] += d[i];
    h[dim] = std::sqrt(d[dim]);
  }
  for (int pass = 0; pass < 2 && !classical; ++pass) {
    GH_TRY(mgs_step(ctx, n, vv, nullptr, 0, mg->mem[0], 8));
    for (int i = 1; i < dim; ++i) GH_TRY(mgs_step(ctx, n, vv, mg->mem[i - 1], 8 + i - 1, mg->mem[i], 8 + i));
    GH_TRY(mgs_step(ctx, n, vv, mg->mem[dim - 1], 8 + dim - 1, nullptr, 8 + dim));
    GH_TRY(finalize_slots(ctx, 8, dim + 1, false));
    GH_TRY(fetch_scalars(ctx, 8, dim + 1, d.data()));
    for (int i = 0; i < dim; ++i) h[i] += d[i];
    h[dim] = std::sqrt(d[dim]);
    // reortho_test (gmres.cc:124-137): sqrt(vv.vv) recomputed by the same
    // deterministic reduction equals h[dim], so the test is h + 1e-3 h == h
    const double oldnorm = h[dim];
    const double newnorm = oldnorm + 1.e-3 * h[dim];
    if (oldnorm != newnorm) break;
  }
  const double s = h[dim];
  GH_TRY(vec_scale_to(ctx, GH_F64, n, vv, 1. / s, vv));
  // givens_rotation (gmres.cc:91-105)
  for (int i = 0; i < last; i++) {
    const double dummy = h[i];
    h[i] = ci[i] * dummy + si[i] * h[i + 1];
    h[i + 1] = -si[i] * dummy + ci[i] * h[i + 1];
  }
  const double r = 1. / std::sqrt(h[last] * h[last] + h[last + 1] * h[last + 1]);
  si[last] = h[last + 1] * r;
  ci[last] = h[last] * r;
  h[last] = ci[last] * h[last] + si[last] * h[last + 1];
  gamma[last + 1] = -si[last] * gamma[last];
  gamma[last] *= ci[last];
  *rho = std::fabs(gamma[dim]);
  return GH_OK;
}

// backward (gmres.cc:235-260) on H1(i,j) = H[j][i]
std::vector<double> gmres_backward(const std::vector<std::vector<double>>& H,
                                   const std::vector<double>& gamma, int dim) {
  std::vector<double> hs(std::max(dim, 1), 0.0);
  for (int i = dim - 1; i >= 0; i--) {
    double s = gamma[i];
    for (int j = i + 1; j < dim; j++) s -= hs[j] * H[j][i];
    hs[i] = s / H[i][i];
  }
  return hs;
}

// GMRES::solve with x_zero: the caller guarantees x == 0 on entry (Newton's
// correction), so residualgmres is p = b with Dirichlet rows zeroed.
int gmres_impl(gh_ctx* ctx, gh_mg* mg, const gh_csr* A, const gh_mask* mask, int coarselevel,
               int vmax, double* x, const double* b, bool x_zero, gh_cginfo* info,
               int* reached_out, int* steps_out) {
  const int64_t n = mg->L[mg->nlevels - 1].size;
  GH_TRY(ensure_gmres_memory(mg, vmax, false));
  std::vector<std::vector<double>> H(vmax);
  for (int i = 0; i < vmax; ++i) H[i].assign(i + 2, 0.0);
  std::vector<double> gamma(vmax + 1, 0.0), ci(vmax, 0.0), si(vmax, 0.0);
  const int k0 = info->iteration;
  int reached = 0;
  double* v = mg->mem[0];
  double* p = mg->p;
  // residualgmres (stdsolver.cc:743-754): p = b - A x; Dirichlet rows zero
  if (x_zero)
    GH_TRY(vec_convert(ctx, n, GH_F64, p, GH_F64, b));
  else
    GH_TRY(csr_residual(ctx, A, GH_F64, p, x, b));
  if (mask) GH_TRY(mask_apply(ctx, mask, GH_F64, p));
  GH_TRY(precond_raw(mg, coarselevel, v, p));
  GH_TRY(reduce_dot(ctx, GH_F64, n, v, v, 2));
  double rr;
  GH_TRY(fetch_scalars(ctx, 2, 1, &rr));
  double rho = std::sqrt(rr);
  int dim = 0;
  if (rho == 0.) {
    *reached_out = 1;
    *steps_out = 0;
    return GH_OK;
  }
  gamma[0] = rho;
  GH_TRY(vec_scale_to(ctx, GH_F64, n, v, 1. / rho, v));
  std::vector<double> d(vmax + 2);
  while (dim < vmax - 1 && !reached) {
    dim++;
    const int last = dim - 1;
    double* vv = mg->mem[dim];
    GH_TRY(csr_spmv(ctx, A, GH_F64, p, GH_F64, mg->mem[last], 1.0, 0.0));  // vmulteq
    GH_TRY(precond_raw(mg, coarselevel, vv, p));
    GH_TRY(arnoldi_step(mg, dim, vv, H[last], gamma, ci, si, d, &rho));
    reached = gh_cginfo_check(info, rho, (double)(k0 + last));
  }
  if (dim > 0) {
    const std::vector<double> hs = gmres_backward(H, gamma, dim);
    GH_TRY(basis_update(mg, x, false, dim, hs.data()));
  }
  // the update stays queued: the caller's next fetched scalar orders it
  *reached_out = reached;
  *steps_out = dim;
  return GH_OK;
}

// Flexible GMRES (right preconditioned; Saad, SIAM J. Sci. Comput. 14 (1993)):
// z_j = M^-1 v_j is kept, the Arnoldi relation is A Z = V H with FP64 products
// A z_j, and x = x0 + Z y.  The preconditioner may then differ from step to
// step -- the lower-precision V-cycle is linear only up to its own rounding --
// and the minimised quantity is the true FP64 residual.  Orthogonalisation,
// re-orthogonalisation test, Givens rotations and back-substitution are those
// of GMRES::solve; its right-preconditioned form (gmres.cc:174-178, 201-205,
// 252-258) applies M once more to V y at the end, which needs a fixed M.
// Start a convergence history at residual res: the first residual and the
// aimed residual res * tol, as gh_cginfo_check records them on its first call.
void cginfo_begin(gh_cginfo* I, double res) {
  I->residual = res;
  I->firstresidual = res;
  I->previousresidual = res;
  I->previouscorrection = 0.;
  I->aimedresidual = res * I->tol;
  I->status = GH_RUNNING;
}

int fgmres_impl(gh_ctx* ctx, gh_mg* mg, const gh_csr* A, const gh_mask* mask, int coarselevel,
                int vmax, double* x, const double* b, bool x_zero, gh_cginfo* info,
                int* reached_out, int* steps_out, double known_rho) {
  const int64_t n = mg->L[mg->nlevels - 1].size;
  GH_TRY(ensure_gmres_memory(mg, vmax, true));
  std::vector<std::vector<double>> H(vmax);
  for (int i = 0; i < vmax; ++i) H[i].assign(i + 2, 0.0);
  std::vector<double> gamma(vmax + 1, 0.0), ci(vmax, 0.0), si(vmax, 0.0);
  const int k0 = info->iteration;
  int reached = 0;
  double* v = mg->mem[0];
  // residualgmres: v = b - A x, Dirichlet rows zero
  if (x_zero)
    GH_TRY(vec_convert(ctx, n, GH_F64, v, GH_F64, b));
  else
    GH_TRY(csr_residual(ctx, A, GH_F64, v, x, b));
  if (mask) GH_TRY(mask_apply(ctx, mask, GH_F64, v));
  double rho = known_rho;
  if (!(x_zero && known_rho >= 0.)) {
    GH_TRY(reduce_dot(ctx, GH_F64, n, v, v, 2));
    double rr;
    GH_TRY(fetch_scalars(ctx, 2, 1, &rr));
    rho = std::sqrt(rr);
  }
  int dim = 0;
  if (rho == 0.) {
    *reached_out = 1;
    *steps_out = 0;
    return GH_OK;
  }
  gamma[0] = rho;
  // The relative tolerance is measured from this starting residual.  (The
  // reference's GMRES, reproduced by gmres_impl, takes the residual after its
  // first step as the base; the flexible variant's estimate is the true FP64
  // residual, so its base is the residual it starts from.)
  cginfo_begin(info, rho);
  GH_TRY(vec_scale_to(ctx, GH_F64, n, v, 1. / rho, v));
  const int fl = mg->nlevels - 1;
  const int clevel = coarsest_level(mg, fl, coarselevel < 0 ? fl : coarselevel);
  std::vector<double> d(vmax + 3);
  while (dim < vmax - 1 && !reached) {
    dim++;
    const int last = dim - 1;
    double* vv = mg->mem[dim];
    void* z = mg->zmem[last];
    // z = M^-1 v_last from a zero start, kept in the V-cycle's precision
    if (mg->P.prec_maxiter != 1) {
      if (mg->wprec != GH_F64)
        return fail(GH_EINVAL, "flexible GMRES with a mixed-precision V-cycle needs prec_maxiter = 1");
      GH_TRY(precond_raw(mg, coarselevel, (double*)z, mg->mem[last]));
    } else {
      GH_TRY(fast_vcycle(mg, clevel, z, mg->wprec, mg->mem[last], GH_F64, true));
    }
    if (mg->wprec == GH_F64)
      GH_TRY(csr_spmv(ctx, A, GH_F64, vv, GH_F64, z, 1.0, 0.0));
    else
      GH_TRY(csr_spmv_x32(ctx, A, vv, (const float*)z));
    GH_TRY(arnoldi_step(mg, dim, vv, H[last], gamma, ci, si, d, &rho, true));
    reached = gh_cginfo_check(info, rho, (double)(k0 + last));
  }
  if (dim > 0) {
    const std::vector<double> hs = gmres_backward(H, gamma, dim);
    GH_TRY(basis_update(mg, x, true, dim, hs.data()));
  }
  *reached_out = reached;
  *steps_out = dim;
  return GH_OK;
}

// known_rho >= 0: the Euclidean norm of b with the mask applied, when the
// caller has it (flexible GMRES from x = 0 then skips that reduction)
int krylov_impl(gh_ctx* ctx, gh_mg* mg, const gh_csr* A, const gh_mask* mask, int coarselevel,
                int vmax, double* x, const double* b, bool x_zero, gh_cginfo* info,
                int* reached_out, int* steps_out, double known_rho = -1.0) {
  if (mg->P.krylov == 1)
    return fgmres_impl(ctx, mg, A, mask, coarselevel, vmax, x, b, x_zero, info, reached_out,
                       steps_out, known_rho);
  return gmres_impl(ctx, mg, A, mask, coarselevel, vmax, x, b, x_zero, info, reached_out,
                    steps_out);
}

double vec_norm_kind(gh_ctx* ctx, const double* x, int64_t n, int kind, int* rc) {
  double out = 0.0;
  if (kind == 0) {
    *rc = reduce_norm_l8(ctx, GH_F64, n, x, 3);
    if (!*rc) *rc = fetch_scalars(ctx, 3, 1, &out);
    return out;
  }
  *rc = reduce_dot(ctx, GH_F64, n, x, x, 3);
  if (!*rc) *rc = fetch_scalars(ctx, 3, 1, &out);
  return std::sqrt(out);
}

// NewtonResidual (stdmultilevelsolver.cc:608-621): y = b - A x, Dirichlet rows 0
int newton_residual(gh_ctx* ctx, const gh_csr* A, const gh_mask* mask, double* y,
                    const double* x, const double* b, int64_t n, int kind, double* nrm) {
  GH_TRY(csr_residual(ctx, A, GH_F64, y, x, b));
  if (mask) GH_TRY(mask_apply(ctx, mask, GH_F64, y));
  int rc = 0;
  *nrm = vec_norm_kind(ctx, y, n, kind, &rc);
  return rc;
}

}  // namespace

extern "C" {

int gh_gmres_solve(gh_ctx* ctx, gh_mg* mg, const gh_csr* A, const gh_mask* mask,
                   int coarselevel, int vmax, gh_vec* x, const gh_vec* b, gh_cginfo* info,
                   int* reached) {
  GH_ARG(ctx && mg && A && x && b && info && vmax >= 2 && vmax <= 256,
         "gh_gmres_solve: bad arguments (2 <= vmax <= 256)");
  const MgLevel& F = mg->L[mg->nlevels - 1];
  GH_ARG(A->prec == GH_F64 && A->n == F.n && A->ncomp == F.ncomp,
         "gh_gmres_solve: A must be the FP64 finest-level matrix");
  GH_ARG(x->prec == GH_F64 && b->prec == GH_F64 && x->n == F.n && b->n == F.n,
         "gh_gmres_solve: x and b must be FP64 finest-level vectors");
  GH_ARG(A == mg->user_fine, "gh_gmres_solve: A must be the finest-level matrix of the solver");
  GH_ARG(!mask || mask == mg->user_mask,
         "gh_gmres_solve: mask must be NULL or the finest-level mask the solver was created with");
  int rdum = 0, steps = 0;
  double* xd = (double*)x->d;
  const double* bd = (const double*)b->d;
  if (F.d_perm) {
    GH_TRY(vec_permute(ctx, F.n, F.ncomp, F.d_perm, GH_F64, mg->ext_x, GH_F64, xd, true));
    GH_TRY(vec_permute(ctx, F.n, F.ncomp, F.d_perm, GH_F64, mg->ext_y, GH_F64, bd, true));
    xd = mg->ext_x;
    bd = mg->ext_y;
  }
  GH_TRY(krylov_impl(ctx, mg, F.A64s ? F.A64s : F.A64, F.M, coarselevel, vmax, xd, bd, false, info,
                     reached ? reached : &rdum, &steps));
  if (F.d_perm)
    GH_TRY(vec_permute(ctx, F.n, F.ncomp, F.d_perm, GH_F64, x->d, GH_F64, mg->ext_x, false));
  GH_HIP(hipStreamSynchronize(ctx->stream));
  return GH_OK;
}

int gh_newton_solve(gh_ctx* ctx, gh_mg* mg, const gh_csr* A, const gh_mask* user_mask,
                    const gh_hn* user_hn, int coarselevel, int vmax, gh_vec* u, const gh_vec* f,
                    gh_cginfo* linfo, gh_nlinfo* nl, int norm_kind, gh_solve_report* rep) {
  GH_ARG(ctx && mg && A && u && f && linfo && nl && vmax >= 2 && vmax <= 256,
         "gh_newton_solve: bad arguments (2 <= vmax <= 256)");
  const MgLevel& F = mg->L[mg->nlevels - 1];
  GH_ARG(A->prec == GH_F64 && A->n == F.n && A->ncomp == F.ncomp,
         "gh_newton_solve: A must be the FP64 finest-level matrix");
  GH_ARG(u->prec == GH_F64 && f->prec == GH_F64 && u->n == F.n && f->n == F.n,
         "gh_newton_solve: u and f must be FP64 finest-level vectors");
  GH_ARG(A == mg->user_fine, "gh_newton_solve: A must be the finest-level matrix of the solver");
  GH_ARG(!user_mask || user_mask == mg->user_mask,
         "gh_newton_solve: mask must be NULL or the finest-level mask the solver was created with");
  GH_ARG(!user_hn || user_hn == mg->user_hn,
         "gh_newton_solve: hn must be NULL or the finest-level hanging nodes the solver was created with");
  const int64_t n = F.size;
  GH_TRY(ensure_gmres_memory(mg, vmax, mg->P.krylov == 1));
  double* r = mg->r;
  double* w = mg->w;
  double* ud = (double*)u->d;
  const double* fd = (const double*)f->d;
  if (F.d_perm) {
    GH_TRY(vec_permute(ctx, F.n, F.ncomp, F.d_perm, GH_F64, mg->ext_x, GH_F64, ud, true));
    GH_TRY(vec_permute(ctx, F.n, F.ncomp, F.d_perm, GH_F64, mg->ext_y, GH_F64, fd, true));
    ud = mg->ext_x;
    fd = mg->ext_y;
  }
  A = F.A64s ? F.A64s : F.A64;
  const gh_mask* mask = F.M;
  const gh_hn* hn = F.H;
  const int64_t mg0 = mg->mgsteps;
  int gm = 0, ndamp = 0, nexpl = 0;
  if (hn) GH_TRY(hn_apply(ctx, hn, u->ncomp, GH_F64, ud, 0));  // Solve: HNAverage(u)
  gh_nlinfo_reset(nl);
  gh_cginfo_reset(linfo);
  double rr;
  {
    // a zero start (the usual case) needs no product: every row of f - A u is
    // then f_i (up to the sign of an exactly zero f_i)
    int rc = 0;
    const double u8 = vec_norm_kind(ctx, ud, n, 0, &rc);
    GH_TRY(rc);
    if (u8 == 0.0) {
      GH_TRY(vec_convert(ctx, n, GH_F64, r, GH_F64, fd));
      if (mask) GH_TRY(mask_apply(ctx, mask, GH_F64, r));
      rr = vec_norm_kind(ctx, r, n, norm_kind, &rc);
      GH_TRY(rc);
    } else {
      GH_TRY(newton_residual(ctx, A, mask, r, ud, fd, n, norm_kind, &rr));
    }
  }
  if (rep) rep->first_residual = rr;
  int reached = gh_nlinfo_check(nl, 0, rr, 0.);
  for (int it = 1; !reached; it++) {
    // NewtonLinearSolve: info.reset(); Zero(x); Gmres
    gh_cginfo_reset(linfo);
    int greached = 0, steps = 0;
    GH_TRY(vec_fill_zero(ctx, GH_F64, n, w));
    // r is masked and, with the Euclidean NewtonNorm, rr is its norm
    GH_TRY(krylov_impl(ctx, mg, A, mask, coarselevel, vmax, w, r, true, linfo, &greached, &steps,
                       norm_kind == 1 ? rr : -1.0));
    gm += steps;
    // NewtonUpdate
    int lex = linfo->status == GH_EXPLODED;
    int rc = 0;
    const double nn = vec_norm_kind(ctx, w, n, norm_kind, &rc);
    GH_TRY(rc);
    // r is unchanged since its norm was taken: with the Euclidean NewtonNorm
    // that norm is the one needed here
    const double nr = norm_kind == 1 ? rr : vec_norm_kind(ctx, r, n, 1, &rc);
    GH_TRY(rc);
    if (nn > 1.e30 || !(nn >= 0.) || nr > 1.e30 || !(nr >= 0.)) lex = 1;
    if (lex) {
      nl->status = GH_NL_DIVERGED;
    } else {
      const double omega = 0.7;
      double relax = 1.;
      GH_TRY(vec_axpby(ctx, GH_F64, n, ud, 1.0, relax, w));
      GH_TRY(newton_residual(ctx, A, mask, r, ud, fd, n, norm_kind, &rr));
      for (int iter = 0; iter < nl->maxrelax; iter++) {
        const int msg = nl_check_damping(nl, iter, rr);
        if (msg == 0) break;
        if (msg == 1) {
          ndamp++;
          GH_TRY(vec_axpby(ctx, GH_F64, n, ud, 1.0, relax * (omega - 1.), w));
          GH_TRY(newton_residual(ctx, A, mask, r, ud, fd, n, norm_kind, &rr));
          relax *= omega;
          continue;
        }
        nexpl = 1;
        GH_TRY(vec_axpby(ctx, GH_F64, n, ud, 1.0, -relax, w));
        relax = 0.;
        nl->status = GH_NL_DIVERGED;
        break;
      }
    }
    reached = gh_nlinfo_check(nl, it, rr, nn);
  }
  if (hn) GH_TRY(hn_apply(ctx, hn, u->ncomp, GH_F64, ud, 2));  // Solve: HNZero(u)
  if (F.d_perm)
    GH_TRY(vec_permute(ctx, F.n, F.ncomp, F.d_perm, GH_F64, u->d, GH_F64, mg->ext_x, false));
  GH_HIP(hipStreamSynchronize(ctx->stream));
  if (rep) {
    rep->newton_iterations = nl->iteration;
    rep->gmres_iterations = gm;
    rep->mgsteps = mg->mgsteps - mg0;
    rep->status = nl->status;
    rep->damping_steps = ndamp;
    rep->damping_exploded = nexpl;
    rep->last_residual = rr;
  }
  return GH_OK;
}

// ------------------------------------------------------------------ A^T x
// One thread per component c of block row i of A^T (block column i of A):
// y(i, c) = d * sum over the blocks (j, i) of A of sum_r A_ji(r, c) x(j, r)
// + b * y(i, c), blocks in ascending j.
__global__ __launch_bounds__(kBlock) void k_csr_t(int64_t n, int B, const int32_t* __restrict__ trow,
                                                  const int32_t* __restrict__ tcol,
                                                  const int32_t* __restrict__ tpos,
                                                  const double* __restrict__ val,
                                                  const double* __restrict__ x,
                                                  double* __restrict__ y, double d, double b,
                                                  int read_y) {
  const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (t >= n * B) return;
  const int64_t i = t / B;
  const int c = (int)(t - i * B);
  const int BB = B * B;
  double s = 0.0;
  for (int k = trow[i]; k < trow[i + 1]; ++k) {
    const double* a = val + (int64_t)tpos[k] * BB + c;
    const double* xj = x + (int64_t)tcol[k] * B;
    for (int r = 0; r < B; ++r) s += a[r * B] * xj[r];
  }
  double v = d * s;
  if (read_y) v += b * y[t];
  y[t] = v;
}

int gh_spmv_t(gh_ctx* ctx, const gh_csr* Ac, gh_vec* y, const gh_vec* x, double d, double b) {
  GH_ARG(ctx && Ac && x && y && x->n == Ac->n && y->n == Ac->n && x->ncomp == Ac->ncomp &&
             y->ncomp == Ac->ncomp,
         "gh_spmv_t: shape mismatch");
  GH_ARG(!Ac->host_only, "gh_spmv_t: the matrix has no device layout (created by gh_csr_create_host)");
  GH_ARG(Ac->prec == GH_F64 && x->prec == GH_F64 && y->prec == GH_F64,
         "gh_spmv_t: FP64 only");
  GH_ARG(x->d != y->d, "gh_spmv_t: x and y must not alias");
  gh_csr* A = const_cast<gh_csr*>(Ac);
  if (!A->t_rowptr) {
    const int64_t n = A->n, nnz = A->nnz;
    std::vector<int32_t> trow(n + 1, 0), tcol(nnz), tpos(nnz);
    for (int64_t k = 0; k < nnz; ++k) trow[A->h_col[k] + 1]++;
    for (int64_t i = 0; i < n; ++i) trow[i + 1] += trow[i];
    std::vector<int32_t> fill(trow.begin(), trow.end() - 1);
    for (int64_t i = 0; i < n; ++i)
      for (int32_t k = A->h_rowptr[i]; k < A->h_rowptr[i + 1]; ++k) {
        const int32_t j = A->h_col[k];
        tcol[fill[j]] = (int32_t)i;
        tpos[fill[j]] = k;
        fill[j]++;
      }
    GH_TRY(dupload(ctx, &A->t_rowptr, trow));
    GH_TRY(dupload(ctx, &A->t_col, tcol));
    GH_TRY(dupload(ctx, &A->t_pos, tpos));
  }
  hipLaunchKernelGGL(k_csr_t, dim3(grid_for(A->n * A->ncomp)), dim3(kBlock), 0, ctx->stream, A->n,
                     A->ncomp, A->t_rowptr, A->t_col, A->t_pos, (const double*)A->val,
                     (const double*)x->d, (double*)y->d, d, b, (int)(b != 0.0));
  GH_HIP(hipGetLastError());
  return GH_OK;
}

}  // extern "C"
