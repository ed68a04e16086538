/*
 * gascoigne_hip.hpp: C++ API surface of the MI355X multigrid solve path.
 *
 * Header-only classes over the C ABI of gascoigne_hip.h that keep the names,
 * argument meaning and call shapes of the reference's cuda/ backend:
 *   CudaVectorInterface      cuda/cudavectorinterface.h:26-77
 *   CudaCSRMatrixInterface   cuda/cudacsrmatrixinterface.h:40-159
 *   CudaMgInterpolatorNested cuda/cudamgintepolatornested.h:17-43
 *   CudaMultiLevelSolver     cuda/cudamultilevelsolver.h:15-45 with the
 *                            StdMultiLevelSolver entry points it drives
 *                            (LinearMg, precondition, Gmres, Solve:
 *                            src/Solver/stdmultilevelsolver.cc:411-458,
 *                            548-758, 940-952, 677-689)
 *   CGInfo / NLInfo          src/Common/cginfo.h:118-146, src/Common/nlinfo.h
 *
 * Errors: the reference terminates the process on a failed CUDA call
 * (cuda/check_cuda.h:28-67).  Here a failed call throws gascoigne_hip::Error
 * carrying the GH_E* code and gh_last_error()'s message; solver outcomes are
 * statuses in CGInfo / NLInfo, as in the reference.
 *
 * Ownership: every object owns its device handle (RAII, move-only).  Objects
 * created on a Context keep the context alive at the C level (reference
 * counted), but must not be used after their Context is destroyed.
 */
#ifndef GASCOIGNE_HIP_HPP
#define GASCOIGNE_HIP_HPP

#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "gascoigne_hip.h"

namespace gascoigne_hip {

class Error : public std::runtime_error {
 public:
  Error(int code, const std::string& msg) : std::runtime_error(msg), code_(code) {}
  int code() const { return code_; }

 private:
  int code_;
};

inline void check(int rc) {
  if (rc != GH_OK) throw Error(rc, std::string("gascoigne_hip: ") + gh_last_error());
}

// Handle holder: move-only, destroys with the C ABI's destroy function.
template <typename T, int (*Destroy)(T*)>
class Handle {
 public:
  Handle() = default;
  explicit Handle(T* p) : p_(p) {}
  Handle(const Handle&) = delete;
  Handle& operator=(const Handle&) = delete;
  Handle(Handle&& o) noexcept : p_(std::exchange(o.p_, nullptr)) {}
  Handle& operator=(Handle&& o) noexcept {
    if (this != &o) {
      reset();
      p_ = std::exchange(o.p_, nullptr);
    }
    return *this;
  }
  ~Handle() { reset(); }
  void reset() {
    if (p_) Destroy(p_);
    p_ = nullptr;
  }
  T* get() const { return p_; }
  T** out() {
    reset();
    return &p_;
  }

 private:
  T* p_ = nullptr;
};

// Device, stream and reduction workspace (the reference keeps cuSPARSE /
// cuBLAS handles per CudaSolver, cuda/cudasolver.h:31-60).
class Context {
 public:
  explicit Context(int device = 0) { check(gh_ctx_create(device, h_.out())); }
  gh_ctx* get() const { return h_.get(); }
  void sync() const { check(gh_ctx_sync(h_.get())); }
  void* stream() const { return gh_ctx_stream(h_.get()); }

 private:
  Handle<gh_ctx, gh_ctx_destroy> h_;
};

// CudaVectorInterface: node-major interleaved GlobalVector on the device.
class CudaVectorInterface {
 public:
  CudaVectorInterface() = default;
  // (n, n_comp): zeroed (cudavectorinterface.cc:28-35)
  CudaVectorInterface(const Context& ctx, int64_t n, int n_comp = 1, int prec = GH_F64)
      : ctx_(&ctx) {
    reserve(n, n_comp, prec);
  }
  // (const GlobalVector&): upload (cudavectorinterface.cc:61-88)
  CudaVectorInterface(const Context& ctx, const std::vector<double>& gv, int n_comp = 1,
                      int prec = GH_F64)
      : ctx_(&ctx) {
    reserve((int64_t)gv.size() / n_comp, n_comp, prec);
    *this = gv;
  }
  CudaVectorInterface& operator=(const std::vector<double>& gv) {
    if ((int64_t)gv.size() != n * n_comp) throw Error(GH_EINVAL, "CudaVectorInterface: size mismatch");
    check(gh_vec_upload(ctx_->get(), h_.get(), gv.data()));
    return *this;
  }
  // operator=(CudaVectorInterface): device copy (converts precision)
  void assign(const CudaVectorInterface& o) { check(gh_vec_copy(ctx_->get(), h_.get(), o.get())); }

  void reserve(int64_t n_, int n_comp_ = 1, int prec = GH_F64) {
    check(gh_vec_create(ctx_->get(), n_, n_comp_, prec, h_.out()));
    n = n_;
    n_comp = n_comp_;
    valuesDevice = gh_vec_data(h_.get());
  }
  void free() {
    h_.reset();
    n = 0;
    valuesDevice = nullptr;
  }
  explicit operator bool() const { return h_.get() != nullptr; }

  void copy_back(std::vector<double>& gv) const {
    gv.resize((size_t)(n * n_comp));
    check(gh_vec_download(ctx_->get(), h_.get(), gv.data()));
  }
  std::vector<double> copy_back() const {
    std::vector<double> gv;
    copy_back(gv);
    return gv;
  }
  void zero() { check(gh_vec_zero(ctx_->get(), h_.get())); }
  void add(double d, const CudaVectorInterface& y) { check(gh_vec_add(ctx_->get(), h_.get(), d, y.get())); }
  void equ(double s, const CudaVectorInterface& y) { check(gh_vec_equ(ctx_->get(), h_.get(), s, y.get())); }
  void scal(double d) { check(gh_vec_scal(ctx_->get(), h_.get(), d)); }
  double norm() const {
    double r;
    check(gh_vec_norm(ctx_->get(), h_.get(), &r));
    return r;
  }
  double dot(const CudaVectorInterface& y) const {
    double r;
    check(gh_vec_dot(ctx_->get(), h_.get(), y.get(), &r));
    return r;
  }

  gh_vec* get() const { return h_.get(); }
  const Context& context() const { return *ctx_; }

  // the reference's public fields
  void* valuesDevice = nullptr;
  int64_t n = 0;
  int n_comp = 1;

 private:
  const Context* ctx_ = nullptr;
  Handle<gh_vec, gh_vec_destroy> h_;
};

// CudaCSRMatrixInterface: node-block CSR (ColumnStencil + FMatrixBlock values)
// on the device, with the inverse block diagonal for Jacobi.
class CudaCSRMatrixInterface {
 public:
  CudaCSRMatrixInterface(const Context& ctx, int64_t n, const std::vector<int32_t>& rowptr,
                         const std::vector<int32_t>& col, const std::vector<double>& val,
                         int n_comp = 1, int prec = GH_F64)
      : ctx_(&ctx) {
    check(gh_csr_create(ctx.get(), n, n_comp, rowptr.data(), col.data(), val.data(), prec,
                        h_.out()));
  }
  // An FP64 level matrix kept on the host only (gh_csr_create_host): for a
  // CudaMultiLevelSolver given an order, which builds its own device matrices;
  // products, residuals and smoothing on it throw.
  struct HostOnly {};
  CudaCSRMatrixInterface(HostOnly, const Context& ctx, int64_t n,
                         const std::vector<int32_t>& rowptr, const std::vector<int32_t>& col,
                         const std::vector<double>& val, int n_comp = 1)
      : ctx_(&ctx) {
    check(gh_csr_create_host(ctx.get(), n, n_comp, rowptr.data(), col.data(), val.data(),
                             h_.out()));
  }
  int64_t n() const { return gh_csr_n(h_.get()); }
  int64_t n_entries() const { return gh_csr_nnz(h_.get()); }
  // y = d*A*x + b*y (cudacsrmatrixinterface.cc:372-412)
  void vmult(CudaVectorInterface& y, const CudaVectorInterface& x, double d = 1.0,
             double b = 0.0) const {
    check(gh_spmv(ctx_->get(), h_.get(), y.get(), x.get(), d, b));
  }
  void vmult_transpose(CudaVectorInterface& y, const CudaVectorInterface& x, double d = 1.0,
                       double b = 0.0) const {
    check(gh_spmv_t(ctx_->get(), h_.get(), y.get(), x.get(), d, b));
  }
  // CudaSolver::MatrixResidual: r = b - A x
  void residual(CudaVectorInterface& r, const CudaVectorInterface& x,
                const CudaVectorInterface& b) const {
    check(gh_residual(ctx_->get(), h_.get(), r.get(), x.get(), b.get()));
  }
  // CudaSolver::smooth: niter sweeps of damped Jacobi (smoother 0) / Richardson (1)
  void smooth(CudaVectorInterface& x, const CudaVectorInterface& b, CudaVectorInterface& h,
              int niter, double omega, int smoother = 0) const {
    check(gh_smooth(ctx_->get(), h_.get(), x.get(), b.get(), h.get(), niter, omega, smoother));
  }
  gh_csr* get() const { return h_.get(); }

 private:
  const Context* ctx_;
  Handle<gh_csr, gh_csr_destroy> h_;
};

// CudaMgInterpolatorNested: BasicInit from the mesh-transfer lists.
class CudaMgInterpolatorNested {
 public:
  // c2f[ncoarse]; zweier rows of 3, vierer rows of 5, achter rows of 9 (fine node, parents)
  CudaMgInterpolatorNested(const Context& ctx, const std::vector<int32_t>& c2f,
                           const std::vector<int32_t>& zweier, const std::vector<int32_t>& vierer,
                           const std::vector<int32_t>& achter)
      : ctx_(&ctx) {
    check(gh_transfer_create(ctx.get(), (int64_t)c2f.size(), c2f.data(), (int64_t)zweier.size() / 3,
                             zweier.data(), (int64_t)vierer.size() / 5, vierer.data(),
                             (int64_t)achter.size() / 9, achter.data(), h_.out()));
  }
  void restrict_zero(CudaVectorInterface& uc, const CudaVectorInterface& uf) const {
    check(gh_restrict_zero(ctx_->get(), h_.get(), uc.get(), uf.get()));
  }
  void prolongate_add(CudaVectorInterface& uf, const CudaVectorInterface& uc) const {
    check(gh_prolongate_add(ctx_->get(), h_.get(), uf.get(), uc.get()));
  }
  gh_transfer* get() const { return h_.get(); }

 private:
  const Context* ctx_;
  Handle<gh_transfer, gh_transfer_destroy> h_;
};

// Hanging-node structure (HNStructureQ12d / Q13d / Q22d / Q23d)
class HNStructure {
 public:
  HNStructure(const Context& ctx, int kind, int64_t n, const std::vector<int32_t>& edges,
              const std::vector<int32_t>& faces)
      : ctx_(&ctx) {
    check(gh_hn_create(ctx.get(), kind, n, (int64_t)edges.size() / 4, edges.data(),
                       (int64_t)faces.size() / 10, faces.data(), h_.out()));
  }
  void Average(CudaVectorInterface& x) const { check(gh_hn_average(ctx_->get(), h_.get(), x.get())); }
  void Distribute(CudaVectorInterface& x) const { check(gh_hn_distribute(ctx_->get(), h_.get(), x.get())); }
  void Zero(CudaVectorInterface& x) const { check(gh_hn_zero(ctx_->get(), h_.get(), x.get())); }
  gh_hn* get() const { return h_.get(); }

 private:
  const Context* ctx_;
  Handle<gh_hn, gh_hn_destroy> h_;
};

// Strong Dirichlet nodes and components (SetBoundaryVectorZero)
class DirichletMask {
 public:
  DirichletMask(const Context& ctx, int64_t n, int ncomp, const std::vector<int32_t>& nodes,
                const std::vector<int32_t>& comps = {0})
      : ctx_(&ctx) {
    check(gh_mask_create(ctx.get(), n, ncomp, (int64_t)nodes.size(), nodes.data(),
                         (int)comps.size(), comps.data(), h_.out()));
  }
  void SetBoundaryVectorZero(CudaVectorInterface& x) const {
    check(gh_mask_zero(ctx_->get(), h_.get(), x.get()));
  }
  gh_mask* get() const { return h_.get(); }

 private:
  const Context* ctx_;
  Handle<gh_mask, gh_mask_destroy> h_;
};

// CGInfo: tolerance, iteration control and status of a linear iteration
struct CGInfo {
  gh_cginfo c{};
  explicit CGInfo(double tol = 1e-6, double globaltol = 1e-14, int maxiter = 100) {
    gh_cginfo_init(&c, tol, globaltol, maxiter);
  }
  void reset() { gh_cginfo_reset(&c); }
  int check(double res, double cor = 0.0) { return gh_cginfo_check(&c, res, cor); }
  std::string status() const {
    static const char* s[] = {"waiting", "running", "converged", "too many iterations",
                              "exploded", "diverging"};
    return c.status >= 0 && c.status <= 5 ? s[c.status] : "unknown";
  }
  int iteration() const { return c.iteration; }
  double residual() const { return c.residual; }
};

// NLInfo: the Newton / iterative-refinement control
struct NLInfo {
  gh_nlinfo c{};
  explicit NLInfo(double tol = 1e-4, double globaltol = 1e-12, int maxiter = 10) {
    gh_nlinfo_init(&c, tol, globaltol, maxiter);
  }
  void reset() { gh_nlinfo_reset(&c); }
  std::string status() const {
    static const char* s[] = {"waiting", "running", "converged", "diverged",
                              "slow_convergence", "stagnation", "exploded"};
    return c.status >= 0 && c.status <= 6 ? s[c.status] : "unknown";
  }
  int iteration() const { return c.iteration; }
};

// Parameters of the MultiLevelSolver / Solver blocks
// (src/Solver/stdmultilevelsolverdata.cc:41-60, src/Solver/solverdata.cc:41-62)
struct MgParams {
  gh_mg_params p{};
  MgParams() { gh_mg_params_default(&p); }
};

// CudaMultiLevelSolver: the whole solve on the device.  Levels in solver
// numbering (0 = coarsest); transfer[l] maps level l to l-1 (transfer[0] may
// be null); hn[l] / mask[l] may be null.  The objects must outlive the solver.
class CudaMultiLevelSolver {
 public:
  CudaMultiLevelSolver(const Context& ctx, const std::vector<const CudaCSRMatrixInterface*>& A,
                       const std::vector<const CudaMgInterpolatorNested*>& transfer,
                       const std::vector<const HNStructure*>& hn,
                       const std::vector<const DirichletMask*>& mask, const MgParams& params,
                       const std::vector<std::vector<int32_t>>* order = nullptr)
      : ctx_(&ctx), nlevels_((int)A.size()) {
    std::vector<gh_csr*> a;
    std::vector<gh_transfer*> t;
    std::vector<gh_hn*> h;
    std::vector<gh_mask*> m;
    for (int l = 0; l < nlevels_; ++l) {
      a.push_back(A[l]->get());
      t.push_back(l < (int)transfer.size() && transfer[l] ? transfer[l]->get() : nullptr);
      h.push_back(l < (int)hn.size() && hn[l] ? hn[l]->get() : nullptr);
      m.push_back(l < (int)mask.size() && mask[l] ? mask[l]->get() : nullptr);
    }
    fine_ = A.back();
    if (order) {
      std::vector<const int32_t*> o;
      for (const auto& v : *order) o.push_back(v.data());
      check(gh_mg_create_ordered(ctx.get(), nlevels_, a.data(), t.data(), h.data(), m.data(),
                                 o.data(), &params.p, h_.out()));
    } else {
      check(gh_mg_create(ctx.get(), nlevels_, a.data(), t.data(), h.data(), m.data(), &params.p,
                         h_.out()));
    }
  }
  int nlevels() const { return nlevels_; }
  int64_t mgsteps() const { return gh_mg_mgsteps(h_.get()); }

  void LinearMg(int finelevel, int coarselevel, CudaVectorInterface& u,
                const CudaVectorInterface& f, CGInfo& info) {
    check(gh_mg_linear(ctx_->get(), h_.get(), finelevel, coarselevel, u.get(), f.get(), &info.c));
  }
  void precondition(CudaVectorInterface& x, const CudaVectorInterface& y, int coarselevel = 0,
                    bool x_is_zero = false) {
    check(gh_mg_precondition(ctx_->get(), h_.get(), coarselevel, x.get(), y.get(), x_is_zero));
  }
  // StdMultiLevelSolver::Gmres: returns the solve's `reached` flag
  int Gmres(CudaVectorInterface& x, const CudaVectorInterface& b, CGInfo& info, int vmax = 10,
            int coarselevel = 0) {
    int reached = 0;
    check(gh_gmres_solve(ctx_->get(), h_.get(), fine_->get(), nullptr, coarselevel, vmax, x.get(),
                         b.get(), &info.c, &reached));
    return reached;
  }
  // StdMultiLevelSolver::Solve (Newton / iterative refinement)
  gh_solve_report Solve(CudaVectorInterface& u, const CudaVectorInterface& f, CGInfo& linfo,
                        NLInfo& nlinfo, int vmax = 10, int norm_kind = 0, int coarselevel = 0) {
    gh_solve_report rep{};
    check(gh_newton_solve(ctx_->get(), h_.get(), fine_->get(), nullptr, nullptr, coarselevel,
                          vmax, u.get(), f.get(), &linfo.c, &nlinfo.c, norm_kind, &rep));
    return rep;
  }
  // Vanka patches of a level (smoother 2): npatches x psize node indices in
  // the caller's numbering (VankaSmoother::ConstructStructure,
  // src/LinAlg/vankasmoother.cc:46-58)
  void SetVankaPatches(int level, int64_t npatches, int psize, const int32_t* nodes) {
    check(gh_mg_set_vanka(h_.get(), level, npatches, psize, nodes));
  }
  // ILU(0) factors of a level (smoother 3), ILU row i = node perm[i] in the
  // caller's numbering, nullptr for the identity (StdSolver::ComputeIlu,
  // src/Solver/stdsolver.cc:1547-1573)
  void SetIluOrder(int level, const int32_t* perm) {
    check(gh_mg_set_ilu(h_.get(), level, perm));
  }
  // The matrix the solver runs on a level (GetSolver(l)->GetMatrix(A)),
  // owned by the solver: FP64, or the V-cycle's precision
  const gh_csr* LevelMatrix(int level, bool fp64 = true) const {
    const gh_csr* A = gh_mg_level_matrix(h_.get(), level, fp64 ? 1 : 0);
    if (!A) throw Error(GH_EINVAL, "LevelMatrix: no such level");
    return A;
  }
  gh_mg* get() const { return h_.get(); }

 private:
  const Context* ctx_;
  int nlevels_;
  const CudaCSRMatrixInterface* fine_ = nullptr;
  Handle<gh_mg, gh_mg_destroy> h_;
};

// VankaSmoother (src/LinAlg/vankasmoother.h:65-125): patch LU factors of an
// FP64 node-block matrix (FP64 or FP32 factors); Solve(x) in place
class VankaSmoother {
 public:
  VankaSmoother(const Context& ctx, const CudaCSRMatrixInterface& A, int64_t npatches, int psize,
                const int32_t* nodes, int prec = GH_F64)
      : ctx_(&ctx) {
    check(gh_vanka_create(ctx.get(), A.get(), npatches, psize, nodes, prec, h_.out()));
  }
  void Solve(CudaVectorInterface& x) const { check(gh_vanka_solve(ctx_->get(), h_.get(), x.get())); }
  gh_vanka* get() const { return h_.get(); }

 private:
  const Context* ctx_;
  Handle<gh_vanka, gh_vanka_destroy> h_;
};

// SparseBlockIlu (src/LinAlg/sparseblockilu.xx): ILU(0) of an FP64 node-block
// matrix (blocks of 1..4) in the order perm (nullptr: identity); Solve(x) in
// place, FP64 or FP32 x
class SparseBlockIlu {
 public:
  SparseBlockIlu(const Context& ctx, const CudaCSRMatrixInterface& A, const int32_t* perm = nullptr)
      : ctx_(&ctx) {
    check(gh_ilu_create(ctx.get(), A.get(), perm, h_.out()));
  }
  void Solve(CudaVectorInterface& x) const { check(gh_ilu_solve(ctx_->get(), h_.get(), x.get())); }
  gh_ilu* get() const { return h_.get(); }

 private:
  const Context* ctx_;
  Handle<gh_ilu, gh_ilu_destroy> h_;
};

}  // namespace gascoigne_hip

#endif
