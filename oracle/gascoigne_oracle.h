/*
 * CPU restatement of Gascoigne's FP64 multigrid / GMRES solve path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path:
 * only the test-suite, the smoke check and the benchmark's CPU-baseline leg
 * load it.  Nothing in the product links or calls it.
 *
 * Every function follows one piece of the reference CPU path and cites it.
 * Vectors are Gascoigne GlobalVectors: node-major, interleaved components,
 * entry (i,c) at i*ncomp+c (src/Common/compvector.h:98-113).
 *
 * Pinning (how each function is checked, see tests/test_oracle.py):
 *   or_vmult / or_residual / or_jacobi / or_norm / or_dot      golden vectors of
 *       the compiled reference (SparseBlockMatrix, nvector) on its own meshes
 *   or_transfer_*                                              golden vectors of
 *       the compiled reference MgInterpolatorMatrix
 *   or_hn_*                                                    golden vectors of
 *       the compiled reference HNStructureQ13d / Q23d
 *   or_dirichlet_zero                                          reference
 *       boundary lists + matrix identity rows
 *   or_cginfo_check / or_nlinfo_check                          hand-traced
 *       status sequences of CGInfo::check / NLInfo::check
 *   or_linear_mg / or_gmres_solve / or_newton_solve            mathematics:
 *       exact two-grid limit, V-cycle error propagation operator vs its
 *       explicit matrix, GMRES minimal-residual property, convergence to the
 *       dense direct solution (the reference's StdMultiLevelSolver/GMRES
 *       compile but cannot link here, see oracle/ref/Makefile)
 *   or_dJdk_cells                                              mathematics
 *       (no reference counterpart): Euler's identity sum_c k_c dJ/dk_c = -J,
 *       the element matrices summing to the assembled matrix, and central
 *       finite differences of J on direct solves
 *   or_block_gauss_jordan                                      inverse times
 *       block = identity, permutation cases
 *   or_direct_*                                                residual of the
 *       solve (the reference's UMFPACK coarse solve is absent: parity with it
 *       is unpinned; only the mathematics pins this function)
 */
#ifndef GASCOIGNE_ORACLE_H
#define GASCOIGNE_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Node-block CSR matrix: n block rows, B = ncomp, val has nnz*B*B entries,
 * each block row-major.  col sorted ascending per row (ColumnStencil). */
typedef struct {
  int n, ncomp;
  const int32_t* rowptr;
  const int32_t* col;
  const double* val;
  int32_t* diag; /* owned, filled by or_matrix_init */
} or_matrix;

int or_matrix_init(or_matrix* A, int n, int ncomp, const int32_t* rowptr,
                   const int32_t* col, const double* val);
void or_matrix_free(or_matrix* A);

void or_vmult(const or_matrix* A, double* y, const double* x, double s);
void or_vmult_transpose(const or_matrix* A, double* y, const double* x, double s);
void or_residual(const or_matrix* A, double* h, const double* x, const double* b);
void or_jacobi(const or_matrix* A, double* h);
void or_block_gauss_jordan(int N, double* m);

double or_norm(const double* x, int64_t n);
double or_norm_l8(const double* x, int64_t n);
double or_dot(const double* x, const double* y, int64_t n);
void or_equ(double* dst, double s, const double* src, int64_t n);
void or_add(double* dst, double s, const double* src, int64_t n);
void or_sadd(double* y, double a, double d, const double* v, int64_t n);
void or_zero(double* x, int64_t n);

/* SparseBlockIlu<FMatrixBlock<B>> (src/LinAlg/sparseblockilu.xx): ILU(0) in
 * the node order perm (ILU row i = node perm[i]; NULL: identity), on A's
 * pattern renumbered with each row's columns ascending.  or_ilu_init runs
 * ConstructStructure (:70-127), copy_entries (:131-166) and compute_ilu
 * (:240-275); or_ilu_solve is solve (:226-236): hin, forward, backward, her.
 * Pinned by the compiled reference's factors and solves (golden fixtures). */
typedef struct {
  int n, ncomp;
  int32_t *p, *q;
  int32_t *rowptr, *col, *diag;
  double* val;
  double* y;
} or_ilu;
int or_ilu_init(or_ilu* I, const or_matrix* A, const int32_t* perm);
void or_ilu_solve(const or_ilu* I, double* x);
void or_ilu_free(or_ilu* I);

/* grid transfer built from c2f / 2er / 4er / 8er maps */
typedef struct {
  int ncoarse, nfine;
  int32_t *crow, *ccol; double* cval; /* restriction, coarse rows */
  int32_t *frow, *fcol; double* fval; /* prolongation, fine rows */
} or_transfer;

or_transfer* or_transfer_create(int ncoarse, const int32_t* c2f, int n2,
                                const int32_t* zweier, int n4, const int32_t* vierer,
                                int n8, const int32_t* achter);
void or_transfer_free(or_transfer* T);
void or_restrict_zero(const or_transfer* T, int ncomp, double* uc, const double* uf);
void or_prolongate_add(const or_transfer* T, int ncomp, double* uf, const double* uc);
int64_t or_transfer_nnz(const or_transfer* T);
void or_transfer_export(const or_transfer* T, int32_t* crow, int32_t* ccol, double* cval,
                        int32_t* frow, int32_t* fcol, double* fval);

/* hanging nodes: kind 12 (Q1 2D), 13 (Q1 3D), 22 (Q2 2D), 23 (Q2 3D).
 * edges: ne x 4 (node, p0, p1, p2); faces: nf x 10 (node, f0..f8),
 * both sorted by node (std::map order). */
typedef struct {
  int kind, ne, nf;
  const int32_t* edges;
  const int32_t* faces;
} or_hn;

void or_hn_average(const or_hn* H, int ncomp, double* u);
void or_hn_distribute(const or_hn* H, int ncomp, double* u);
void or_hn_zero(const or_hn* H, int ncomp, double* u);

/* strong Dirichlet zeroing of listed nodes / components */
void or_dirichlet_zero(int nidx, const int32_t* idx, int ncomp, int ncomps,
                       const int32_t* comps, double* u);

/* CGInfo (src/Common/cginfo.h:118-146) */
enum { OR_WAITING = 0, OR_RUNNING, OR_CONVERGED, OR_TOO_MANY, OR_EXPLODED, OR_DIVERGING };
typedef struct {
  double tol, globaltol, breaktol;
  int miniter, maxiter;
  int status, iteration, totaliter;
  double firstresidual, residual, aimedresidual, previousresidual;
  double correction, previouscorrection, rate, lastrate;
} or_cginfo;

void or_cginfo_init(or_cginfo* I, double tol, double globaltol, int maxiter);
void or_cginfo_reset(or_cginfo* I);
int or_cginfo_check(or_cginfo* I, double res, double cor);

/* NLInfo (src/Common/nlinfo.h) */
enum { OR_NL_WAITING = 0, OR_NL_RUNNING, OR_NL_CONVERGED, OR_NL_DIVERGED,
       OR_NL_SLOW, OR_NL_STAGNATION, OR_NL_EXPLODED };
typedef struct {
  double tol, globaltol, breaktol, rho, linrho, maxresincrease;
  int miniter, maxiter, maxrelax;
  int status, iteration, laststepbad, relax, totaliter;
  double firstresidual, residual, aimedresidual, previousresidual;
  double correction, previouscorrection, rate, lastrate;
} or_nlinfo;

void or_nlinfo_init(or_nlinfo* I, double tol, double globaltol, int maxiter);
void or_nlinfo_reset(or_nlinfo* I);
int or_nlinfo_check(or_nlinfo* I, int iter, double res, double cor);
int or_nlinfo_check_damping(or_nlinfo* I, int dampit, double res); /* 0 ok,1 continue,2 exploded */

/* One multigrid level (solver numbering: 0 = coarsest). */
typedef struct {
  or_matrix A;
  or_transfer* T;   /* transfer from this level to level-1 (NULL on level 0) */
  or_hn hn;
  int ndir;         /* Dirichlet node list */
  const int32_t* dir;
  int ndircomp;
  const int32_t* dircomp;
  int direct;       /* 1: dense direct coarse solve (nnodes < ndirect) */
  double* lu;       /* dense LU of A (direct levels) */
  int32_t* piv;
  void* vanka;      /* or_vanka* of the level (smoother 3), not owned */
  void* ilu;        /* or_ilu* of the level (smoother 4), not owned */
} or_level;

typedef struct {
  int nlevels;
  or_level* L;
  char mgtype;      /* 'V','W','F' */
  int iterpre, iterpost, iterexact;
  double omega, mgomega;
  int smoother;     /* 0 jacobi, 1 richardson, 2 none, 3 vanka ("ilu" = VankaSmoother),
                       4 ilu (SparseBlockIlu) */
  int prec_maxiter;
  double prec_tol, prec_globaltol;
  int64_t mgsteps;  /* counter ("--> mgstep") */
  double **u, **b, **v; /* per level work vectors */
  /* planted mistakes, for tests that show a check fails on them (0 in every
   * other use): OR_PLANT_* bits drop or reorder one step of mgstep */
  int plant;
} or_mg;

enum {
  OR_PLANT_NO_HN_DISTRIBUTE = 1, /* restriction without HNDistribute */
  OR_PLANT_NO_HN_AVERAGE = 2,    /* prolongation without HNAverage */
  OR_PLANT_NO_HN_ZERO_FINE = 4   /* prolonged correction without HNZero(v) */
};

or_mg* or_mg_create(int nlevels);
void or_mg_free(or_mg* M);
int or_mg_setup(or_mg* M); /* allocates work vectors, factors direct levels */
void or_linear_mg(or_mg* M, int finelevel, int coarselevel, double* u,
                  const double* f, or_cginfo* info);
void or_precondition(or_mg* M, int coarselevel, double* x, const double* y);

/* GMRES(vmax) with left preconditioning (src/LinAlg/gmres.cc) */
int or_gmres_solve(or_mg* M, int coarselevel, int vmax, double* x, const double* b,
                   or_cginfo* info);

/* Newton / iterative refinement driver (StdMultiLevelSolver::newton).
 * norm_kind 0: NewtonNorm = max norm (reference); 1: Euclidean norm. */
typedef struct {
  int newton_iterations, gmres_iterations;
  int64_t mgsteps;
  int status;
  double first_residual, last_residual;
  int damping_steps;    /* "continue" answers of check_damping, all steps */
  int damping_exploded; /* 1 when check_damping answered "exploded" */
} or_solve_report;

int or_newton_solve(or_mg* M, int coarselevel, int vmax, double* u, const double* f,
                    or_cginfo* linfo, or_nlinfo* nlinfo, int norm_kind,
                    or_solve_report* rep);

/* Vanka smoother (src/LinAlg/vankasmoother.cc): patches of `psize` nodes,
 * nodes[p*psize + r] (local order = local unknown order, components
 * interleaved); each patch matrix A_p = A restricted to the patch's unknowns is
 * LU-factored with partial pivoting (Eigen PartialPivLU in the reference,
 * :111, absent here: an unblocked LU with the same pivot rule, first largest
 * |a_ik| in column k).  solve: x <- (1/count_i) sum_p R_p^T A_p^-1 R_p x, the
 * patches added in ascending order (:193-237). */
typedef struct {
  int n, ncomp, npatches, psize, m;
  int32_t* nodes;
  double* lu;   /* npatches * m * m, row-major, L unit lower below the diagonal */
  int32_t* piv; /* npatches * m: row swapped with row k at step k */
} or_vanka;

or_vanka* or_vanka_create(const or_matrix* A, int npatches, int psize, const int32_t* nodes);
void or_vanka_free(or_vanka* V);
void or_vanka_solve(const or_vanka* V, double* x);

/* Coefficient gradient of J(u(k)) for -div(k grad u) = f with one coefficient
 * per cell of an axis-aligned Q1 box mesh (no reference counterpart: the
 * reference has no adjoint solve).  g[c] = -sum_{a,b} z(v_a) K^c_ab u(v_b),
 * K^c the unit-coefficient element matrix (tensor products of the 1D Q1
 * stiffness and mass on the cell's edge lengths), cells in Gascoigne's corner
 * order, xyz[n*3] node coordinates. */
void or_dJdk_cells(int dim, const double* xyz, int64_t ncells, const int32_t* cells,
                   const double* z, const double* u, double* g);

/* OpenMP threads used by the loops the reference parallelises */
int or_num_threads(void);
/* timing variant: 1 also runs the vector operations and the Jacobi
 * application on all threads (norm and dot sums then change order) */
void or_set_parallel_vectors(int on);

#ifdef __cplusplus
}
#endif
#endif
