/*
 * gascoigne_hip: C ABI of the MI355X (gfx950) multigrid solve path.
 *
 * Drop-in for the device side of Gascoigne's cuda/ backend.  Every entry point
 * names the reference interface it replaces (paths relative to the reference
 * repository).  Conventions:
 *   - every call returns GH_OK (0) or a negative GH_E* code; gh_last_error()
 *     gives the message of the last failure on the calling thread.  Solver
 *     outcomes are reported through gh_cginfo / gh_nlinfo status fields, as
 *     CGInfo / NLInfo strings are in the reference (src/Common/cginfo.cc:202-256,
 *     src/Common/nlinfo.cc:247-356).  Nothing calls exit().
 *   - host arrays are borrowed for the duration of the call (copied on create);
 *     device memory belongs to the handle that allocated it.
 *   - vectors are Gascoigne GlobalVectors: node-major, interleaved components,
 *     entry (i,c) at i*ncomp+c (src/Common/compvector.h:98-113).
 *   - indices are int32 (the reference's device CSR is int32 too,
 *     cuda/cudacsrmatrixinterface.h:40-159).
 *   - all work is enqueued on the context's HIP stream; only calls that return
 *     host scalars (norms, dots, solver drivers) synchronise.
 */
#ifndef GASCOIGNE_HIP_H
#define GASCOIGNE_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  GH_OK = 0,
  GH_EINVAL = -1,   /* invalid argument / shape mismatch */
  GH_EHIP = -2,     /* HIP runtime error */
  GH_ENOMEM = -3,   /* device allocation failed */
  GH_ESINGULAR = -4 /* singular diagonal block or coarse matrix */
};

/* storage precision of vectors and matrix values */
enum { GH_F64 = 0, GH_F32 = 1, GH_F16 = 2 };

typedef struct gh_ctx gh_ctx;
typedef struct gh_vec gh_vec;
typedef struct gh_csr gh_csr;
typedef struct gh_transfer gh_transfer;
typedef struct gh_hn gh_hn;
typedef struct gh_mask gh_mask;
typedef struct gh_mg gh_mg;
typedef struct gh_cells gh_cells;
typedef struct gh_patches gh_patches;
typedef struct gh_vanka gh_vanka;
typedef struct gh_ilu gh_ilu;

const char* gh_last_error(void);
const char* gh_version(void);

/* ---------------------------------------------------------------- context */
/* Replaces the per-solver cuSPARSE/cuBLAS handles and the legacy default
 * stream of cuda/cudasolver.h:31-60.  device < 0 keeps the current device. */
int gh_ctx_create(int device, gh_ctx** out);
int gh_ctx_destroy(gh_ctx* ctx);
int gh_ctx_sync(gh_ctx* ctx);
/* the hipStream_t all work of this context is enqueued on */
void* gh_ctx_stream(gh_ctx* ctx);

/* ---------------------------------------------------------------- vectors */
/* CudaVectorInterface (cuda/cudavectorinterface.h:26-77, .cc:16-218):
 * reserve (:141-157) zeroes, like the (n, n_comp) constructor. */
int gh_vec_create(gh_ctx* ctx, int64_t n, int ncomp, int prec, gh_vec** out);
int gh_vec_destroy(gh_vec* v);
int64_t gh_vec_n(const gh_vec* v);
int gh_vec_ncomp(const gh_vec* v);
int gh_vec_prec(const gh_vec* v);
void* gh_vec_data(gh_vec* v);
/* H2D from a GlobalVector (operator=(GlobalVector), :61-88) */
int gh_vec_upload(gh_ctx* ctx, gh_vec* v, const double* host);
/* D2H (copy_back, :110-117); synchronises */
int gh_vec_download(gh_ctx* ctx, const gh_vec* v, double* host);
/* zero (:119-126) */
int gh_vec_zero(gh_ctx* ctx, gh_vec* v);
/* D2D copy (operator=(CudaVectorInterface), :91-103); converts precision */
int gh_vec_copy(gh_ctx* ctx, gh_vec* dst, const gh_vec* src);
/* dst = s*src: CudaSolver::Equ (cuda/cudasolver.cc:403-415) */
int gh_vec_equ(gh_ctx* ctx, gh_vec* dst, double s, const gh_vec* src);
/* dst += s*src: add = cublasDaxpy (:162-174), CudaSolver::Add (:417-429) */
int gh_vec_add(gh_ctx* ctx, gh_vec* dst, double s, const gh_vec* src);
/* y = a*y + d*v: nvector::sadd (src/Common/nvector.h:507-517) */
int gh_vec_sadd(gh_ctx* ctx, gh_vec* y, double a, double d, const gh_vec* v);
/* x *= s: scal = cublasDscal (:185-190) */
int gh_vec_scal(gh_ctx* ctx, gh_vec* x, double s);
/* Euclidean norm: norm = cublasDnrm2 (:176-183), CudaSolver::Norm (:431-447) */
int gh_vec_norm(gh_ctx* ctx, const gh_vec* x, double* out);
/* scalar product: StdSolver::ScalarProduct (src/Solver/stdsolver.cc:1807-1810) */
int gh_vec_dot(gh_ctx* ctx, const gh_vec* x, const gh_vec* y, double* out);
/* max norm: StdSolver::NewtonNorm = GlobalVector::norm_l8 (stdsolver.cc:693-697) */
int gh_vec_norm_l8(gh_ctx* ctx, const gh_vec* x, double* out);

/* ---------------------------------------------------------------- matrices */
/* CudaCSRMatrixInterface(handle, const MatrixInterface&)
 * (cuda/cudacsrmatrixinterface.cc:143-286) + get_inverse_diagonal (:119-141,
 * 291-352, cuda/cusparsehelper.cu:24-118).  Node-block CSR of n block rows
 * with B = ncomp: rowptr[n+1], col[nnz] (sorted, ColumnStencil layout),
 * val[nnz*B*B] row-major blocks.  The matrix is stored in `prec` (GH_F64,
 * GH_F32, or GH_F16 values with FP32 arithmetic); the inverse block diagonal
 * (Gauss-Jordan with partial pivoting, src/LinAlg/nodematrix.h:159-213) is
 * computed in FP64 and kept in FP64 (prec GH_F64) or FP32. */
int gh_csr_create(gh_ctx* ctx, int64_t n, int ncomp, const int32_t* rowptr,
                  const int32_t* col, const double* val, int prec, gh_csr** out);
/* An FP64 level matrix kept on the host only (same arguments and checks as
 * gh_csr_create, nothing allocated on the device; no reference counterpart).
 * gh_mg_create_ordered builds its own device matrices from the host copies,
 * so a caller handing it such matrices skips the device layout of the
 * caller-numbered levels; the finest one is then also the handle its
 * gh_gmres_solve / gh_newton_solve take.  Every device operation (products,
 * residuals, smoothing, gh_csr_layout, gh_vanka_create, gh_mg_create) refuses
 * it with GH_EINVAL. */
int gh_csr_create_host(gh_ctx* ctx, int64_t n, int ncomp, const int32_t* rowptr,
                       const int32_t* col, const double* val, gh_csr** out);
int gh_csr_destroy(gh_csr* A);
int64_t gh_csr_n(const gh_csr* A);
int64_t gh_csr_nnz(const gh_csr* A);
int gh_csr_prec(const gh_csr* A);
/* Device layout (no reference counterpart; for byte accounting), out[8].
 * Scalar matrices: out[0] entries stored (exact zeros of the pattern are not
 * stored; every row sum takes its other terms in the same order, so the values
 * of A x are unchanged), out[1] padded slots, out[2] slices of 64 rows,
 * out[3] slices with 16-bit column offsets, out[4] slices whose rows all follow
 * shared relative column patterns (no column storage), out[5] entries of the
 * pattern table, out[6] column bytes one sweep reads, out[7] 1 when the rows
 * run in slices in another order than their own (rows of one length, and of
 * one pattern, grouped inside windows of 4096 rows), else 0.  Block matrices
 * with 2..4 components (stored as component rows, sorted by length inside
 * windows): out[0] block rows kept (row r of a block is left out when it is
 * all zero), out[1] padded slots, out[2] slices, out[3..5] = 0, out[6] =
 * 4 out[0], out[7] = 1; other block sizes: out[0] = out[1] = nnz blocks,
 * out[2..5] = 0, out[6] = 4 nnz, out[7] = 0. */
int gh_csr_layout(const gh_csr* A, int64_t* out);
/* The values the device holds, in A's CSR order (nnz*B*B, reduced-precision
 * values widened to FP64); the D2H half of operator<< (print of the device
 * matrix, cuda/cudacsrmatrixinterface.cc:439).  Synchronises. */
int gh_csr_download_values(gh_ctx* ctx, const gh_csr* A, double* val);
/* y = d*A*x + b*y: CudaCSRMatrixInterface::vmult (cudacsrmatrixinterface.cc:372-412) */
int gh_spmv(gh_ctx* ctx, const gh_csr* A, gh_vec* y, const gh_vec* x, double d, double b);
/* y = d*A^T*x + b*y, FP64; block matrices are transposed blockwise
 * (adjoint path; no reference counterpart on the device:
 * SparseBlockMatrix::transpose / AssembleDualMatrix, src/Solver/stdsolver.cc:1815-1834) */
int gh_spmv_t(gh_ctx* ctx, const gh_csr* A, gh_vec* y, const gh_vec* x, double d, double b);
/* r = b - A*x: CudaSolver::MatrixResidual (cuda/cudasolver.cc:379-398) */
int gh_residual(gh_ctx* ctx, const gh_csr* A, gh_vec* r, const gh_vec* x, const gh_vec* b);
/* CudaSolver::smooth (cuda/cudasolver.cc:344-374) / StdSolver::smooth
 * (src/Solver/stdsolver.cc:951-986): niter sweeps of
 *   smoother 0 (jacobi):     x += omega * D^{-1} (b - A x)
 *   smoother 1 (richardson): x += omega * (b - A x)
 * h is scratch of x's shape; on return it holds no defined value. */
int gh_smooth(gh_ctx* ctx, const gh_csr* A, gh_vec* x, const gh_vec* b, gh_vec* h,
              int niter, double omega, int smoother);
/* h = D^{-1} h: SparseBlockMatrix::Jacobi (src/LinAlg/sparseblockmatrix.xx:47-56).
 * Allocates (and frees) a temporary of h's size on every call. */
int gh_jacobi(gh_ctx* ctx, const gh_csr* A, gh_vec* h);
/* gh_jacobi without any allocation; no counterpart in the reference.  For
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  The temporary is a
 * fixed 8 MiB device array per GPU, held until the process ends: h may hold at
 * most 8 MiB (1,048,576 FP64 or 2,097,152 FP32 entries), larger h is GH_EINVAL.
 * Every call on one GPU shares that array, so calls that overlap on one GPU
 * (several contexts or host threads) race; calls on one context's stream are
 * ordered.  Nothing synchronises; the result equals gh_jacobi's bit for bit. */
int gh_jacobi_noalloc(gh_ctx* ctx, const gh_csr* A, gh_vec* h);

/* ----------------------------------------------------------------- Vanka */
/* VankaSmoother (src/LinAlg/vankasmoother.h:65-125, .cc:7-237): patches of
 * `psize` nodes, nodes[p*psize + r] (local unknown order: node r, components
 * interleaved; at most 128 unknowns per patch).  ConstructStructure +
 * copy_entries (:7-111): the block of the FP64 matrix A on each patch's
 * unknowns, LU with partial pivoting (first largest |a_ik|), factors stored in
 * `prec` (GH_F64 or GH_F32).  Every node must lie in a patch. */
int gh_vanka_create(gh_ctx* ctx, const gh_csr* A, int64_t npatches, int psize,
                    const int32_t* nodes, int prec, gh_vanka** out);
int gh_vanka_destroy(gh_vanka* V);
/* solve (:193-237), in place: x <- (1/count_i) sum_p R_p^T A_p^-1 R_p x, the
 * patches of a node summed in ascending order.  x in the factors' precision. */
int gh_vanka_solve(gh_ctx* ctx, const gh_vanka* V, gh_vec* x);

/* ------------------------------------------------------------------ ILU(0) */
/* SparseBlockIlu (src/LinAlg/sparseblockilu.xx), the CPU path's default
 * smoother "ilu" (src/Solver/solverdata.cc:55).  ConstructStructure +
 * copy_entries (:70-166): ILU row i is node perm[i] (NULL: the identity), its
 * columns the matrix row's columns in the inverse order, sorted ascending;
 * compute_ilu (:240-275) on the device, rows of one level of the lower part's
 * dependency graph at a time, each row's operations in the reference's order
 * (the factors equal the CPU factors bit for bit); the diagonal blocks hold
 * their inverses.  A is FP64 with blocks of 1..4 components and its values
 * from creation (not device-assembled).  GH_ESINGULAR on a zero pivot. */
int gh_ilu_create(gh_ctx* ctx, const gh_csr* A, const int32_t* perm, gh_ilu** out);
int gh_ilu_destroy(gh_ilu* I);
/* solve (:226-236), in place: x <- (LU)^-1 x through the order perm, FP64 x
 * bit for bit as the CPU ILU, FP32 x on an FP32 copy of the factors.  The
 * triangular solves run level by level. */
int gh_ilu_solve(gh_ctx* ctx, const gh_ilu* I, gh_vec* x);
/* Level counts of the forward (lower) and backward (upper) schedules. */
int gh_ilu_levels(const gh_ilu* I, int32_t* forward, int32_t* backward);
/* The factors in ILU numbering (n+1 rowptr, nnz col, n diagonal positions,
 * nnz*B*B values, row-major blocks); any pointer may be NULL. */
int gh_ilu_download(gh_ctx* ctx, const gh_ilu* I, int32_t* rowptr, int32_t* col, int32_t* diag,
                    double* val);

/* ---------------------------------------------------------------- transfer */
/* CudaMgInterpolatorNested::BasicInit / check_matrices
 * (cuda/cudamgintepolatornested.cc:27-110) and the CPU MgInterpolatorMatrix
 * (src/Discretization/Q1/mginterpolatormatrix.cc:34-121): restriction and
 * prolongation from the mesh-transfer lists c2f[ncoarse], zweier[n2][3],
 * vierer[n4][5], achter[n8][9] (fine node, then parents). */
int gh_transfer_create(gh_ctx* ctx, int64_t ncoarse, const int32_t* c2f, int64_t n2,
                       const int32_t* zweier, int64_t n4, const int32_t* vierer,
                       int64_t n8, const int32_t* achter, gh_transfer** out);
int gh_transfer_destroy(gh_transfer* T);
int64_t gh_transfer_nfine(const gh_transfer* T);
/* uc = R uf: restrict_zero (cudamgintepolatornested.cc:131-139) */
int gh_restrict_zero(gh_ctx* ctx, const gh_transfer* T, gh_vec* uc, const gh_vec* uf);
/* uf += P uc: prolongate_add (cudamgintepolatornested.cc:160-167) */
int gh_prolongate_add(gh_ctx* ctx, const gh_transfer* T, gh_vec* uf, const gh_vec* uc);

/* ---------------------------------------------------------- hanging nodes */
/* CudaSolver::HNZero/HNDistribute/HNAverage (cuda/cudasolver.cc:449-489) with
 * the CPU weights (the device path of the reference selects weights by a
 * discretisation name that never matches, cudasolver.cc:188-237):
 * kind 12 HNStructureQ12d, 13 HNStructureQ13d, 22 HNStructureQ22d,
 * 23 HNStructureQ23d.  edges[ne][4] = (node, p0, p1, p2),
 * faces[nf][10] = (node, f0..f8), each sorted by node (std::map order). */
int gh_hn_create(gh_ctx* ctx, int kind, int64_t n, int64_t ne, const int32_t* edges,
                 int64_t nf, const int32_t* faces, gh_hn** out);
int gh_hn_destroy(gh_hn* H);
int gh_hn_average(gh_ctx* ctx, const gh_hn* H, gh_vec* x);
int gh_hn_distribute(gh_ctx* ctx, const gh_hn* H, gh_vec* x);
int gh_hn_zero(gh_ctx* ctx, const gh_hn* H, gh_vec* x);

/* -------------------------------------------------------------- Dirichlet */
/* CudaSolver::SetProblem Dirichlet mask (cuda/cudasolver.cc:272-297) and
 * SetBoundaryVectorZero (:491-501); CPU CGDisc::StrongDirichletVectorZero
 * (src/DofHandler/cgdisc.h:1162-1179). */
int gh_mask_create(gh_ctx* ctx, int64_t n, int ncomp, int64_t nidx, const int32_t* idx,
                   int ncomps, const int32_t* comps, gh_mask** out);
int gh_mask_destroy(gh_mask* M);
int gh_mask_zero(gh_ctx* ctx, const gh_mask* M, gh_vec* x);

/* ------------------------------------------------------------ solver info */
/* CGInfo (src/Common/cginfo.h:118-146) */
enum { GH_WAITING = 0, GH_RUNNING, GH_CONVERGED, GH_TOO_MANY, GH_EXPLODED, GH_DIVERGING };
typedef struct {
  double tol, globaltol, breaktol;
  int miniter, maxiter;
  int status, iteration, totaliter;
  double firstresidual, residual, aimedresidual, previousresidual;
  double correction, previouscorrection, rate, lastrate;
} gh_cginfo;
void gh_cginfo_init(gh_cginfo* I, double tol, double globaltol, int maxiter);
void gh_cginfo_reset(gh_cginfo* I);
int gh_cginfo_check(gh_cginfo* I, double res, double cor);

/* NLInfo (src/Common/nlinfo.h) */
enum { GH_NL_WAITING = 0, GH_NL_RUNNING, GH_NL_CONVERGED, GH_NL_DIVERGED, GH_NL_SLOW,
       GH_NL_STAGNATION, GH_NL_EXPLODED };
typedef struct {
  double tol, globaltol, breaktol, rho, linrho, maxresincrease;
  int miniter, maxiter, maxrelax;
  int status, iteration, laststepbad, relax, totaliter;
  double firstresidual, residual, aimedresidual, previousresidual;
  double correction, previouscorrection, rate, lastrate;
} gh_nlinfo;
void gh_nlinfo_init(gh_nlinfo* I, double tol, double globaltol, int maxiter);
void gh_nlinfo_reset(gh_nlinfo* I);
int gh_nlinfo_check(gh_nlinfo* I, int iter, double res, double cor);

/* ------------------------------------------------------------- multigrid */
/* Parameters of src/Solver/stdmultilevelsolverdata.cc:41-60 and
 * src/Solver/solverdata.cc:41-62. */
typedef struct {
  char mgtype;          /* 'V', 'W', 'F' */
  int iterpre, iterpost, iterexact;
  double omega, mgomega;
  int smoother;         /* 0 jacobi, 1 richardson, 2 vanka ("ilu" with the Vanka
                           patch smoother: h = b - A x; Vanka solve of h;
                           x += omega h; patches from gh_mg_set_vanka),
                           3 ilu (the same with SparseBlockIlu, the reference's
                           default smoother; factors from gh_mg_set_ilu) */
  int ndirect;          /* levels with fewer nodes get a dense FP64 direct solve */
  int prec_maxiter;
  double prec_tol, prec_globaltol;
  int prec;             /* GH_F64: plain FP64 V-cycle; GH_F32 / GH_F16: FP32
                           arithmetic with FP32 / FP16-stored matrix values */
  int use_graph;        /* 1: replay the V-cycle from a captured HIP graph */
  int krylov;           /* outer Krylov method of gh_gmres_solve / gh_newton_solve:
                           0: GMRES left-preconditioned (gmres.cc:160-233);
                           1: flexible GMRES, right-preconditioned with the
                           preconditioned directions kept, so that a V-cycle
                           in lower precision may serve as the preconditioner
                           while the Arnoldi process minimises the FP64
                           residual (no reference counterpart).  Its CGInfo
                           tolerance is relative to the residual it starts
                           from; GMRES, like the reference's, measures it
                           from the residual after its first step */
  int adjoint;          /* 1: the solver works on the transposed level matrices:
                           gh_gmres_solve / gh_newton_solve / gh_mg_linear then
                           solve A^T z = g for the finest-level A given at
                           creation, with smoothing, residuals and products by
                           A_l^T and restriction = transpose of prolongation (no
                           reference counterpart: the reference has no adjoint
                           solve, src/Solver/stdsolver.cc:1068-1090).  With
                           hanging nodes the level matrices are condensed
                           (hanging rows and columns identity-only); their
                           transposes are condensed over the same constraints,
                           so the V-cycle applies the same HNAverage,
                           HNDistribute and HNZero as the forward solver. */
} gh_mg_params;
void gh_mg_params_default(gh_mg_params* p);

/* CudaMultiLevelSolver (cuda/cudamultilevelsolver.h:15-45): NewSolver
 * (:70-74, one level per CudaSolver) + NewMgInterpolator (:76-104).  Levels in
 * solver numbering (0 = coarsest).  A[l] FP64 matrices (the V-cycle makes its
 * own copies in params->prec), T[l] the transfer from level l to l-1 (T[0]
 * unused), H[l] hanging nodes or NULL, M[l] Dirichlet masks or NULL.
 * The handles must outlive the gh_mg. */
int gh_mg_create(gh_ctx* ctx, int nlevels, gh_csr* const* A, gh_transfer* const* T,
                 gh_hn* const* H, gh_mask* const* M, const gh_mg_params* params,
                 gh_mg** out);
/* As gh_mg_create, with the solver renumbering every level for locality:
 * order[l][k] is the caller's node placed at internal position k (a
 * permutation of 0..n_l-1, e.g. the nodes sorted lexicographically by their
 * coordinates).  The caller's handles are read once and copied in the new
 * numbering; each matrix / transfer row keeps its entries in the caller's
 * order, so every sum is formed as without renumbering.  Vectors passed to
 * the solver stay in the caller's numbering: they are permuted on entry and
 * exit.  No counterpart in the reference, whose device path runs in
 * Gascoigne's hash-ordered numbering (src/Mesh/levelmesh3d.cc:83-94). */
int gh_mg_create_ordered(gh_ctx* ctx, int nlevels, gh_csr* const* A, gh_transfer* const* T,
                         gh_hn* const* H, gh_mask* const* M, const int32_t* const* order,
                         const gh_mg_params* params, gh_mg** out);
/* The Vanka patches of one level (in the caller's node numbering), for
 * smoother 2; built on the solver's FP64 level matrix, factors in the V-cycle's
 * precision.  Levels solved directly need none. */
int gh_mg_set_vanka(gh_mg* mg, int level, int64_t npatches, int psize, const int32_t* nodes);
/* The ILU(0) factors of one level for smoother 3, ILU row i = the caller's
 * node perm[i] (NULL: the caller's numbering), built on the solver's FP64 level
 * matrix (StdSolver::ComputeIlu, src/Solver/stdsolver.cc:1547-1573, without
 * the optional modify_ilu).  Levels solved directly need none. */
int gh_mg_set_ilu(gh_mg* mg, int level, const int32_t* perm);
int gh_mg_destroy(gh_mg* mg);
int gh_mg_nlevels(const gh_mg* mg);
int64_t gh_mg_mgsteps(const gh_mg* mg);  /* the "--> mgstep" counter */
/* The matrix level l of the solver runs, in its internal numbering (the
 * caller's when it was created unordered): fp64 != 0 the FP64 operator of
 * the outer iteration and the FP64 V-cycle, fp64 == 0 the V-cycle's operator
 * in the solver's precision.  Owned by the solver (valid until
 * gh_mg_destroy; never pass it to gh_csr_destroy).  Counterpart:
 * StdMultiLevelSolver::GetSolver(l)->GetMatrix(A)
 * (src/Solver/stdmultilevelsolver.h:150-158, src/Solver/stdsolver.h:335-342).
 * NULL for a bad level. */
const gh_csr* gh_mg_level_matrix(const gh_mg* mg, int level, int fp64);

/* LinearMg (cuda/cudamultilevelsolver.cc:106-165,
 * src/Solver/stdmultilevelsolver.cc:411-458): MG iteration on finelevel from
 * the content of u, until info stops it. */
int gh_mg_linear(gh_ctx* ctx, gh_mg* mg, int finelevel, int coarselevel, gh_vec* u,
                 const gh_vec* f, gh_cginfo* info);
/* StdMultiLevelSolver::precondition (stdmultilevelsolver.cc:940-952):
 * LinearMg on the finest level with the prec_* CGInfo.  x holds the initial
 * guess (x_is_zero = 1 lets the first smoothing sweep skip A*x). */
int gh_mg_precondition(gh_ctx* ctx, gh_mg* mg, int coarselevel, gh_vec* x,
                       const gh_vec* y, int x_is_zero);

/* GMRES::solve (src/LinAlg/gmres.cc:160-260), left preconditioned by
 * gh_mg_precondition, FP64 Arnoldi (modified Gram-Schmidt with the reference's
 * re-orthogonalisation test), Givens rotations and back-substitution.
 * A must be the finest-level matrix the solver was created with; the solve
 * uses that level's Dirichlet mask and hanging nodes given at creation; mask
 * must be NULL or that same handle (anything else is GH_EINVAL).
 * Returns the solve's `reached` flag in *reached. */
int gh_gmres_solve(gh_ctx* ctx, gh_mg* mg, const gh_csr* A, const gh_mask* mask,
                   int coarselevel, int vmax, gh_vec* x, const gh_vec* b, gh_cginfo* info,
                   int* reached);

/* StdMultiLevelSolver::Solve / newton / NewtonLinearSolve / NewtonUpdate
 * (stdmultilevelsolver.cc:548-570, 646-673, 693-758, 832-856): FP64 iterative
 * refinement whose correction is one GMRES(vmax) cycle.  The element residual
 * `Form` is the assembled matrix residual f - A u with Dirichlet rows zeroed.
 * A must be the solver's finest-level matrix; mask and hn are the solver's
 * finest-level handles given at creation: each argument must be NULL or that
 * same handle (anything else is GH_EINVAL).
 * norm_kind 0: NewtonNorm is the max norm (reference); 1: Euclidean norm. */
typedef struct {
  int newton_iterations, gmres_iterations;
  int64_t mgsteps;
  int status;
  double first_residual, last_residual;
  /* NewtonUpdate damping (stdmultilevelsolver.cc:735-752): "continue" answers
   * of NLInfo::check_damping over all steps, and 1 when one answered
   * "exploded" (the step was taken back and the status set to diverged) */
  int damping_steps, damping_exploded;
} gh_solve_report;
int gh_newton_solve(gh_ctx* ctx, gh_mg* mg, const gh_csr* A, const gh_mask* mask,
                    const gh_hn* hn, int coarselevel, int vmax, gh_vec* u, const gh_vec* f,
                    gh_cginfo* linfo, gh_nlinfo* nlinfo, int norm_kind,
                    gh_solve_report* rep);

/* --------------------------------------------------------------- adjoint */
/* Coefficient gradients of J(u(k)) for -div(k grad u) = f assembled with one
 * coefficient per cell (A = sum_c k_c K^c, Dirichlet rows and columns
 * eliminated).  No reference counterpart: the reference assembles dual
 * matrices (StdSolver::AssembleDualMatrix, src/Solver/stdsolver.cc:1815-1834)
 * but has no adjoint solve or coefficient gradient.
 *
 * gh_cells_create: the Q1 cells of an axis-aligned box mesh, ncells x (4 | 8)
 * node indices in Gascoigne's corner order (hex: (0,0,0) (1,0,0) (1,1,0)
 * (0,1,0) then the same at z = 1), node coordinates xyz[n*3]. */
int gh_cells_create(gh_ctx* ctx, int dim, int64_t n, const double* xyz, int64_t ncells,
                    const int32_t* cells, gh_cells** out);
int gh_cells_destroy(gh_cells* C);
/* g[c] = dJ/dk_c = -z^T (dA/dk_c) u = -sum_{a,b} z(v_a) K^c_ab u(v_b), with K^c
 * the unit-coefficient Q1 Laplace element matrix of cell c (tensor products of
 * the 1D stiffness and mass matrices, as the assembly forms it), u the state
 * and z the adjoint (A^T z = dJ/du).  u and z must vanish on Dirichlet nodes
 * (then only free-node couplings contribute, as in the eliminated matrix). */
int gh_dJdk_cells(gh_ctx* ctx, const gh_cells* C, const gh_vec* z, const gh_vec* u, gh_vec* g);

/* ------------------------------------------------- DWR error indicators */
/* The dual-weighted residual estimator of a Q1 solution with the patch-wise
 * Q2 interpolation of the adjoint: what StdLoop::Estimator
 * (src/Solver/stdloop.cc:221-273) would compute from the adjoint solve; the
 * reference holds only stubs there (every estimator type aborts).
 *
 * gh_patches_create: the patches of a refined mesh (the 2^dim children of one
 * coarser cell), npatches x 3^dim node indices in lexicographic order (x
 * fastest), node coordinates xyz[nnodes*3]; every patch must be an
 * axis-aligned box whose middle nodes are midpoints (GH_EINVAL otherwise).
 *
 * gh_dwr_estimate: per cell K (npatches x 2^dim values, each patch's cells in
 * lexicographic order),
 *   eta_K = int_K f w - k_K int_K grad u_h . grad w,   w = I2 z_h - z_h,
 * u_h, z_h the Q1 functions of the nodal state u and adjoint z (A^T z = g for
 * J(u) = g^T u), I2 the Q2 interpolation on the patch, f the Q2 interpolant of
 * the nodal data f (the load b_i = int f phi_i), k the per-cell coefficient
 * (NULL: 1), all scalar FP64.  sum_K eta_K estimates J(u) - J(u_h).  With
 * hanging nodes pass u and z after HNAverage. */
int gh_patches_create(gh_ctx* ctx, int dim, int64_t nnodes, const double* xyz, int64_t npatches,
                      const int32_t* nodes, gh_patches** out);
int gh_patches_destroy(gh_patches* P);
int gh_dwr_estimate(gh_ctx* ctx, const gh_patches* P, const gh_vec* u, const gh_vec* z,
                    const gh_vec* f, const gh_vec* k, gh_vec* eta);

/* ------------------------------------------------------ device assembly */
/* CudaSolver::AssembleMatrix (cuda/cudasolver.cc:114-128): StdSolver's CPU
 * assembly CGDisc::Matrix (src/DofHandler/cgdisc.h:599-635, element matrices of
 * src/DofHandler/elementintegrator.xx:350-397) with SparseBlockMatrix::dirichlet
 * (src/LinAlg/sparseblockmatrix.xx:205-218), uploaded and its diagonal inverted
 * (get_inverse_diagonal, cuda/cudacsrmatrixinterface.cc:119-141) -- here formed
 * on the device.  The Q1 operator of -div(k grad u) on the box cells C with one
 * coefficient per cell (k: FP64 vector of C's cell count, NULL: k = 1) is
 * written into the existing scalar matrix A (any precision, any layout; its
 * pattern must hold every coupling of the cells), with the rows and columns of
 * the mask's nodes made identity, and A's inverse diagonal re-formed.  Each
 * entry sums its cells' contributions in ascending cell order, so the values
 * equal a host assembly in that order bit for bit.  An entry that was an exact
 * zero when A was created (left out of the device layout) must stay zero, else
 * GH_EINVAL.  Afterwards A's creation values are gone: gh_mg_create of a
 * non-FP64, ordered or adjoint solver refuses A.  Synchronises. */
int gh_assemble_q1(gh_ctx* ctx, const gh_cells* C, const gh_vec* k, const gh_mask* dirichlet,
                   gh_csr* A);
/* CGDisc::Form (src/DofHandler/cgdisc.h:415-448) as used by
 * StdMultiLevelSolver::NewtonResidual (src/Solver/stdmultilevelsolver.cc:608-621):
 * r = f - A(u) formed cell by cell from u (no matrix), each node summing its
 * cells in ascending order; the mask's rows set to zero (SetBoundaryVectorZero).
 * u enters at every node, Dirichlet nodes included. */
int gh_form_q1(gh_ctx* ctx, const gh_cells* C, const gh_vec* k, const gh_mask* dirichlet,
               const gh_vec* u, const gh_vec* f, gh_vec* r);
/* The same for the triquadratic patch elements of CGDiscQ23d
 * (src/DofHandler/cgdisc.h:1292; the Q2 Lagrange basis of BaseQ23d on the 27
 * nodes of a patch, src/Discretization/Q2/baseq23d.h): the operator of
 * -div(k grad u) with one coefficient per patch element (k: FP64 vector of
 * P's patch count, NULL: k = 1) on the patches P of gh_patches_create (dim 3),
 * written into the scalar matrix A like gh_assemble_q1.  Each entry sums its
 * patches' contributions in ascending patch order (the host producer's
 * order), and A's columns must ascend in every row (GH_EINVAL otherwise).
 * Unconstrained: on meshes with hanging nodes the condensation is the
 * caller's.  Synchronises. */
int gh_assemble_q2(gh_ctx* ctx, const gh_patches* P, const gh_vec* k, const gh_mask* dirichlet,
                   gh_csr* A);
/* The Q2 operator condensed over the hanging nodes (CGDisc::LocalToGlobal
 * with HNStructureQ23d::CondenseHanging, src/DofHandler/cgdisc.h:353-366,
 * src/Discretization/Q2/hnstructureq23d.cc:120-156, in global form):
 * Ac = T^T A T from the unconstrained FP64 operator A (gh_assemble_q2), T the
 * interpolation of the edge constraints (nedges x [h, near, middle, far],
 * weights 3/8, 3/4, -1/8) and the face constraints (nfaces x [h, 9 parents],
 * products of those weights); hanging rows are identity rows, then the
 * mask's rows and columns.  Ac: a scalar matrix of the condensed pattern with
 * ascending columns (any precision), its values written like
 * gh_assemble_q2's; the values equal the host producer's condensation (the
 * sign of a zero sum aside).  Synchronises. */
int gh_condense_q2(gh_ctx* ctx, int64_t nedges, const int32_t* edges, int64_t nfaces,
                   const int32_t* faces, const gh_csr* A, const gh_mask* dirichlet, gh_csr* Ac);
/* CGDisc::Form (src/DofHandler/cgdisc.h:415-448) for the same elements:
 * r = f - A(u) patch by patch from u, each node summing its patches in
 * ascending order; the mask's rows set to zero. */
int gh_form_q2(gh_ctx* ctx, const gh_patches* P, const gh_vec* k, const gh_mask* dirichlet,
               const gh_vec* u, const gh_vec* f, gh_vec* r);
/* StdMultiLevelSolver::AssembleMatrix + ComputeIlu
 * (src/Solver/stdmultilevelsolver.cc:763-790): gh_assemble_q1 on every level
 * of a solver, into the level operators it runs on (the FP64 operator, the
 * reduced-precision V-cycle copy and the flexible-GMRES copy; for a solver
 * made by gh_mg_create those FP64 operators are the caller's matrices),
 * followed by the inverse diagonals and, for directly solved levels, the dense
 * inverse (formed on the host from the new values).  C[l] and k[l] (k or k[l]
 * NULL: k = 1) in the caller's numbering of level l; the solver's own Dirichlet
 * masks apply.  Levels with hanging nodes, the Vanka smoother and adjoint
 * solvers are refused (GH_EINVAL).  Synchronises. */
int gh_mg_assemble_q1(gh_ctx* ctx, gh_mg* mg, gh_cells* const* C, gh_vec* const* k);
/* The same for a hierarchy of triquadratic patch elements (gh_assemble_q2 on
 * every level; P[l] and k[l], one coefficient per patch, in the caller's
 * numbering of level l).  Levels with hanging nodes (whose operators are
 * condensed), the Vanka and ILU smoothers and adjoint solvers are refused
 * (GH_EINVAL).  Synchronises. */
int gh_mg_assemble_q2(gh_ctx* ctx, gh_mg* mg, gh_patches* const* P, gh_vec* const* k);

#ifdef __cplusplus
}
#endif
#endif
