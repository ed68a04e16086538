/*
 * CPU restatement of Gascoigne's FP64 multigrid / GMRES solve path.
 * TEST INFRASTRUCTURE ONLY (the checker for the HIP path); see the header.
 *
 * Arithmetic follows the reference expression by expression (operation order
 * and association), so that the compiled reference's golden vectors are met
 * bit for bit where the reference is deterministic.  Build without FMA
 * contraction (-ffp-contract=off), as the reference's x86-64 build has none.
 */
#include "gascoigne_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* ---------------------------------------------------------------- Vanka */
/* copy_entries_sparseblockmatrix (vankasmoother.cc:64-111): the block of A on
 * the patch's unknowns, then PartialPivLU::compute */
or_vanka* or_vanka_create(const or_matrix* A, int npatches, int psize, const int32_t* nodes) {
  const int B = A->ncomp, m = psize * B;
  or_vanka* V = (or_vanka*)calloc(1, sizeof(or_vanka));
  V->n = A->n;
  V->ncomp = B;
  V->npatches = npatches;
  V->psize = psize;
  V->m = m;
  V->nodes = (int32_t*)malloc(sizeof(int32_t) * (size_t)npatches * psize);
  memcpy(V->nodes, nodes, sizeof(int32_t) * (size_t)npatches * psize);
  V->lu = (double*)calloc((size_t)npatches * m * m, sizeof(double));
  V->piv = (int32_t*)calloc((size_t)npatches * m, sizeof(int32_t));
#pragma omp parallel for schedule(static)
  for (int p = 0; p < npatches; ++p) {
    const int32_t* pn = nodes + (int64_t)p * psize;
    double* a = V->lu + (int64_t)p * m * m;
    int32_t* pv = V->piv + (int64_t)p * m;
    for (int r = 0; r < psize; ++r) {
      const int row = pn[r];
      for (int pos = A->rowptr[row]; pos < A->rowptr[row + 1]; ++pos) {
        const int col = A->col[pos];
        int c = -1;
        for (int q = 0; q < psize; ++q)
          if (pn[q] == col) c = q;
        if (c < 0) continue;
        for (int cr = 0; cr < B; ++cr)
          for (int cc = 0; cc < B; ++cc)
            a[(B * r + cr) * m + B * c + cc] = A->val[((int64_t)pos * B + cr) * B + cc];
      }
    }
    for (int k = 0; k < m; ++k) {
      int pr = k;
      double best = fabs(a[k * m + k]);
      for (int i = k + 1; i < m; ++i)
        if (fabs(a[i * m + k]) > best) {
          best = fabs(a[i * m + k]);
          pr = i;
        }
      pv[k] = pr;
      if (pr != k)
        for (int j = 0; j < m; ++j) {
          const double t = a[k * m + j];
          a[k * m + j] = a[pr * m + j];
          a[pr * m + j] = t;
        }
      const double d = a[k * m + k];
      if (d == 0.0) continue;
      for (int i = k + 1; i < m; ++i) {
        const double l = a[i * m + k] / d;
        a[i * m + k] = l;
        for (int j = k + 1; j < m; ++j) a[i * m + j] -= l * a[k * m + j];
      }
    }
  }
  return V;
}

void or_vanka_free(or_vanka* V) {
  if (!V) return;
  free(V->nodes);
  free(V->lu);
  free(V->piv);
  free(V);
}

/* VankaSmoother::solve (vankasmoother.cc:193-237) */
void or_vanka_solve(const or_vanka* V, double* x) {
  const int B = V->ncomp, m = V->m, n = V->n;
  double* b = (double*)malloc(sizeof(double) * (size_t)n * B);
  int* count = (int*)calloc((size_t)n, sizeof(int));
  double* h = (double*)malloc(sizeof(double) * m);
  memcpy(b, x, sizeof(double) * (size_t)n * B);
  memset(x, 0, sizeof(double) * (size_t)n * B);
  for (int p = 0; p < V->npatches; ++p) {
    const int32_t* pn = V->nodes + (int64_t)p * V->psize;
    const double* a = V->lu + (int64_t)p * m * m;
    const int32_t* pv = V->piv + (int64_t)p * m;
    for (int r = 0; r < V->psize; ++r)
      for (int c = 0; c < B; ++c) h[B * r + c] = b[(int64_t)pn[r] * B + c];
    for (int k = 0; k < m; ++k)
      if (pv[k] != k) {
        const double t = h[k];
        h[k] = h[pv[k]];
        h[pv[k]] = t;
      }
    /* triangular solves, column by column (Eigen's blocked order is not
       available to pin; this order is the device's) */
    for (int k = 0; k < m; ++k)
      for (int i = k + 1; i < m; ++i) h[i] -= a[i * m + k] * h[k];
    for (int k = m - 1; k >= 0; --k) {
      h[k] /= a[k * m + k];
      for (int i = 0; i < k; ++i) h[i] -= a[i * m + k] * h[k];
    }
    for (int r = 0; r < V->psize; ++r) {
      count[pn[r]]++;
      for (int c = 0; c < B; ++c) x[(int64_t)pn[r] * B + c] += h[B * r + c];
    }
  }
  for (int i = 0; i < n; ++i) {
    const double s = 1.0 / count[i];
    for (int c = 0; c < B; ++c) x[(int64_t)i * B + c] *= s;
  }
  free(b);
  free(count);
  free(h);
}

/* ------------------------------------------------------------ adjoint */
static double or_s1(int p, int q, double h) { return (p == q ? 1.0 : -1.0) / h; }
static double or_m1(int p, int q, double h) { return h * (p == q ? 2.0 : 1.0) / 6.0; }

void or_dJdk_cells(int dim, const double* xyz, int64_t ncells, const int32_t* cells,
                   const double* z, const double* u, double* g) {
  static const int corner[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                                   {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
  const int npc = dim == 3 ? 8 : 4;
  for (int64_t c = 0; c < ncells; ++c) {
    const int32_t* v = cells + c * npc;
    const double* x0 = xyz + 3 * (int64_t)v[0];
    const double hx = xyz[3 * (int64_t)v[1]] - x0[0];
    const double hy = xyz[3 * (int64_t)v[3] + 1] - x0[1];
    const double hz = dim == 3 ? xyz[3 * (int64_t)v[4] + 2] - x0[2] : 1.0;
    double s = 0.0;
    for (int a = 0; a < npc; ++a)
      for (int b = 0; b < npc; ++b) {
        const int* pa = corner[a];
        const int* pb = corner[b];
        double e;
        if (dim == 3)
          e = or_s1(pa[0], pb[0], hx) * or_m1(pa[1], pb[1], hy) * or_m1(pa[2], pb[2], hz) +
              or_m1(pa[0], pb[0], hx) * or_s1(pa[1], pb[1], hy) * or_m1(pa[2], pb[2], hz) +
              or_m1(pa[0], pb[0], hx) * or_m1(pa[1], pb[1], hy) * or_s1(pa[2], pb[2], hz);
        else
          e = or_s1(pa[0], pb[0], hx) * or_m1(pa[1], pb[1], hy) +
              or_m1(pa[0], pb[0], hx) * or_s1(pa[1], pb[1], hy);
        s += z[v[a]] * e * u[v[b]];
      }
    g[c] = -s;
  }
}

/* 0: the reference's threading (OpenMP over the matrix-vector products and
 * grid transfers only); 1: also the vector operations and the Jacobi
 * application (a timing variant: the sums of norms and dots then run in
 * another order, so iterates differ from the reference's by rounding). */
static int g_parvec = 0;
void or_set_parallel_vectors(int on) { g_parvec = on; }

int or_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* ------------------------------------------------------------------ matrix */

/* ColumnDiagStencil::memory: diag(i) = first position with col == i
 * (src/LinAlg/columndiagstencil.cc:31-46). */
int or_matrix_init(or_matrix* A, int n, int ncomp, const int32_t* rowptr,
                   const int32_t* col, const double* val) {
  A->n = n;
  A->ncomp = ncomp;
  A->rowptr = rowptr;
  A->col = col;
  A->val = val;
  A->diag = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
  if (!A->diag) return -1;
  for (int i = 0; i < n; ++i) {
    A->diag[i] = -1;
    for (int pos = rowptr[i]; pos < rowptr[i + 1]; ++pos)
      if (col[pos] == i) {
        A->diag[i] = pos;
        break;
      }
    if (A->diag[i] < 0) return -2;
  }
  return 0;
}

void or_matrix_free(or_matrix* A) {
  free(A->diag);
  A->diag = NULL;
}

/* FMatrixBlock<N>::cadd (src/LinAlg/fmatrixblock.h:110-122):
 * for each row p: sum = sum_q m(p,q) x(q) (sequential from 0.); y(p) += s*sum. */
static inline void block_cadd(int B, double s, double* p, const double* pm,
                              const double* q0) {
  for (int r = 0; r < B; ++r) {
    double sum = 0.;
    for (int c = 0; c < B; ++c) sum += *pm++ * q0[c];
    p[r] += s * sum;
  }
}

/* SparseBlockMatrix<B>::vmult (src/LinAlg/sparseblockmatrix.xx:378-392),
 * OpenMP over rows as in the reference. */
void or_vmult(const or_matrix* A, double* y, const double* x, double s) {
  const int B = A->ncomp, BB = B * B;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < A->n; ++i) {
    double* off = y + (size_t)i * B;
    for (int pos = A->rowptr[i]; pos < A->rowptr[i + 1]; ++pos)
      block_cadd(B, s, off, A->val + (size_t)pos * BB, x + (size_t)A->col[pos] * B);
  }
}

/* y += s * A^T x (FMatrixBlock::caddtrans order per block,
 * src/LinAlg/fmatrixblock.h:123-135); serial: the scatter has races. */
void or_vmult_transpose(const or_matrix* A, double* y, const double* x, double s) {
  const int B = A->ncomp, BB = B * B;
  for (int i = 0; i < A->n; ++i) {
    const double* xi = x + (size_t)i * B;
    for (int pos = A->rowptr[i]; pos < A->rowptr[i + 1]; ++pos) {
      const double* m = A->val + (size_t)pos * BB;
      double* yj = y + (size_t)A->col[pos] * B;
      for (int r = 0; r < B; ++r)
        for (int c = 0; c < B; ++c) yj[c] += s * (m[r * B + c] * xi[r]);
    }
  }
}

/* StdSolver::MatrixResidual (src/Solver/stdsolver.cc:783-791):
 * Equ(h,1.,b); vmult(A,h,x,-1.). */
void or_residual(const or_matrix* A, double* h, const double* x, const double* b) {
  or_equ(h, 1., b, (int64_t)A->n * A->ncomp);
  or_vmult(A, h, x, -1.);
}

/* NodeMatrix<N>::gauss_jordan (src/LinAlg/nodematrix.h:159-213), m row-major. */
void or_block_gauss_jordan(int N, double* m) {
  int p[64];
  double hv[64];
#define V(i, j) m[(i) * N + (j)]
  for (int i = 0; i < N; i++) p[i] = i;
  for (int j = 0; j < N; j++) {
    double max = fabs(V(j, j));
    int r = j;
    for (int i = j + 1; i < N; i++)
      if (fabs(V(i, j)) > max) {
        max = fabs(V(i, j));
        r = i;
      }
    if (r > j) {
      for (int k = 0; k < N; k++) {
        double hr = V(j, k);
        V(j, k) = V(r, k);
        V(r, k) = hr;
      }
      int t = p[j];
      p[j] = p[r];
      p[r] = t;
    }
    double hr = 1. / V(j, j);
    V(j, j) = hr;
    for (int k = 0; k < N; k++) {
      if (k == j) continue;
      for (int i = 0; i < N; i++) {
        if (i == j) continue;
        V(i, k) -= V(i, j) * V(j, k) * hr;
      }
    }
    for (int i = 0; i < N; i++) {
      V(i, j) *= hr;
      V(j, i) *= -hr;
    }
    V(j, j) = hr;
  }
  for (int i = 0; i < N; i++) {
    for (int k = 0; k < N; k++) hv[p[k]] = V(i, k);
    for (int k = 0; k < N; k++) V(i, k) = hv[k];
  }
#undef V
}

/* SparseBlockMatrix<B>::Jacobi (src/LinAlg/sparseblockmatrix.xx:47-56):
 * Ad = diag block; Ad.inverse(); Ad.vmult(x_i)  (fmatrixblock.h:194-213). */
void or_jacobi(const or_matrix* A, double* h) {
  const int B = A->ncomp, BB = B * B;
#pragma omp parallel for schedule(static) if (g_parvec)
  for (int i = 0; i < A->n; ++i) {
    double Ad[64 * 64], vh[64];
    memcpy(Ad, A->val + (size_t)A->diag[i] * BB, sizeof(double) * BB);
    or_block_gauss_jordan(B, Ad);
    double* p = h + (size_t)i * B;
    for (int c = 0; c < B; ++c) vh[c] = p[c];
    const double* q = Ad;
    for (int r = 0; r < B; ++r) {
      p[r] = 0.;
      for (int c = 0; c < B; ++c) p[r] += *q++ * vh[c];
    }
  }
}

/* ------------------------------------------------------------------ ILU(0) */

static const int32_t* g_sort_key;
static int key_cmp(const void* a, const void* b) {
  const int32_t x = g_sort_key[*(const int32_t*)a], y = g_sort_key[*(const int32_t*)b];
  return (x > y) - (x < y);
}

int or_ilu_init(or_ilu* I, const or_matrix* A, const int32_t* perm) {
  const int n = A->n, B = A->ncomp, BB = B * B;
  const int64_t nnz = A->rowptr[n];
  memset(I, 0, sizeof(*I));
  I->n = n;
  I->ncomp = B;
  I->p = (int32_t*)malloc(sizeof(int32_t) * (n ? n : 1));
  I->q = (int32_t*)malloc(sizeof(int32_t) * (n ? n : 1));
  I->rowptr = (int32_t*)malloc(sizeof(int32_t) * (n + 1));
  I->col = (int32_t*)malloc(sizeof(int32_t) * (nnz ? nnz : 1));
  I->diag = (int32_t*)malloc(sizeof(int32_t) * (n ? n : 1));
  I->val = (double*)calloc((size_t)(nnz ? nnz : 1) * BB, sizeof(double));
  I->y = (double*)malloc(sizeof(double) * (size_t)(n ? n : 1) * B);
  /* ConstructStructure: p = perm, q its inverse; row i lists q[col] of row
   * p[i], sorted ascending; diag = position of column i */
  for (int i = 0; i < n; ++i) I->p[i] = perm ? perm[i] : i;
  for (int i = 0; i < n; ++i) I->q[I->p[i]] = i;
  int32_t zmax = 1;
  for (int i = 0; i < n; ++i) {
    const int32_t r = A->rowptr[i + 1] - A->rowptr[i];
    if (r > zmax) zmax = r;
  }
  int32_t* picol = (int32_t*)malloc(sizeof(int32_t) * zmax);
  int32_t* ppi = (int32_t*)malloc(sizeof(int32_t) * zmax);
  I->rowptr[0] = 0;
  for (int i = 0; i < n; ++i) {
    const int pi = I->p[i];
    const int ni = A->rowptr[pi + 1] - A->rowptr[pi];
    I->rowptr[i + 1] = I->rowptr[i] + ni;
    for (int c = 0; c < ni; ++c) picol[c] = I->q[A->col[A->rowptr[pi] + c]];
    for (int c = 0; c < ni; ++c) ppi[c] = c;
    g_sort_key = picol;
    qsort(ppi, (size_t)ni, sizeof(int32_t), key_cmp);
    for (int c = 0; c < ni; ++c) I->col[I->rowptr[i] + c] = picol[ppi[c]];
    I->diag[i] = -1;
    for (int32_t pos = I->rowptr[i]; pos < I->rowptr[i + 1]; ++pos)
      if (I->col[pos] == i) I->diag[i] = pos;
  }
  free(picol);
  free(ppi);
  /* copy_entries: entry (p[i], j) of A adds into (i, q[j]) */
  for (int i = 0; i < n; ++i) {
    const int pi = I->p[i];
    for (int32_t posA = A->rowptr[pi]; posA < A->rowptr[pi + 1]; ++posA) {
      const int pj = I->q[A->col[posA]];
      for (int32_t pos = I->rowptr[i]; pos < I->rowptr[i + 1]; ++pos)
        if (I->col[pos] == pj) {
          for (int e = 0; e < BB; ++e) I->val[(size_t)pos * BB + e] += A->val[(size_t)posA * BB + e];
          break;
        }
    }
  }
  /* compute_ilu: a(i,k) *= a(k,k)^-1 (operator*=), a(i,j) -= a(i,k) a(k,j)
   * (submult), then a(i,i) <- a(i,i)^-1 (gauss_jordan) */
  double vhelp[8];
  for (int i = 0; i < n; ++i) {
    if (I->diag[i] < 0) return -1;
    for (int32_t pk = I->rowptr[i]; pk < I->diag[i]; ++pk) {
      const int k = I->col[pk];
      double* M = I->val + (size_t)pk * BB;
      const double* D = I->val + (size_t)I->diag[k] * BB;
      for (int r = 0; r < B; ++r) {
        for (int c = 0; c < B; ++c) vhelp[c] = M[r * B + c];
        for (int c = 0; c < B; ++c) {
          M[r * B + c] = 0.;
          for (int t = 0; t < B; ++t) M[r * B + c] += vhelp[t] * D[t * B + c];
        }
      }
      for (int32_t pj = I->diag[k] + 1; pj < I->rowptr[k + 1]; ++pj) {
        const int j = I->col[pj];
        for (int32_t ph = I->rowptr[i]; ph < I->rowptr[i + 1]; ++ph) {
          if (I->col[ph] == j) {
            double* H = I->val + (size_t)ph * BB;
            const double* C = I->val + (size_t)pj * BB;
            for (int r = 0; r < B; ++r)
              for (int c = 0; c < B; ++c)
                for (int t = 0; t < B; ++t) H[r * B + c] -= M[r * B + t] * C[t * B + c];
            break;
          }
        }
      }
    }
    or_block_gauss_jordan(B, I->val + (size_t)I->diag[i] * BB);
  }
  return 0;
}

void or_ilu_solve(const or_ilu* I, double* x) {
  const int n = I->n, B = I->ncomp, BB = B * B;
  double* y = I->y;
  double vh[8];
  for (int i = 0; i < n; ++i)
    for (int c = 0; c < B; ++c) y[(size_t)i * B + c] = 1. * x[(size_t)I->p[i] * B + c];
  /* forward (:193-203) */
  for (int i = 1; i < n; ++i)
    for (int32_t pos = I->rowptr[i]; pos < I->diag[i]; ++pos) {
      const double* M = I->val + (size_t)pos * BB;
      double* yi = y + (size_t)i * B;
      const double* yc = y + (size_t)I->col[pos] * B;
      for (int r = 0; r < B; ++r)
        for (int c = 0; c < B; ++c) yi[r] -= M[r * B + c] * yc[c];
    }
  /* backward (:207-222) */
  for (int i = n - 1; i >= 0; --i) {
    double* yi = y + (size_t)i * B;
    for (int32_t pos = I->rowptr[i + 1] - 1; pos > I->diag[i]; --pos) {
      const double* M = I->val + (size_t)pos * BB;
      const double* yc = y + (size_t)I->col[pos] * B;
      for (int r = 0; r < B; ++r)
        for (int c = 0; c < B; ++c) yi[r] -= M[r * B + c] * yc[c];
    }
    const double* D = I->val + (size_t)I->diag[i] * BB;
    for (int c = 0; c < B; ++c) vh[c] = yi[c];
    for (int r = 0; r < B; ++r) {
      yi[r] = 0.;
      for (int c = 0; c < B; ++c) yi[r] += D[r * B + c] * vh[c];
    }
  }
  for (int i = 0; i < n; ++i)
    for (int c = 0; c < B; ++c) x[(size_t)i * B + c] = 1. * y[(size_t)I->q[i] * B + c];
}

void or_ilu_free(or_ilu* I) {
  free(I->p);
  free(I->q);
  free(I->rowptr);
  free(I->col);
  free(I->diag);
  free(I->val);
  free(I->y);
  memset(I, 0, sizeof(*I));
}

/* ------------------------------------------------------------- nvector ops */
/* src/Common/nvector.h:169-180 (norm), 332-342 (operator*), 365-374 (equ),
 * 451-460 (add), 507-517 (sadd); serial as in the reference. */
double or_norm(const double* x, int64_t n) {
  double s = 0.;
#pragma omp parallel for schedule(static) reduction(+ : s) if (g_parvec)
  for (int64_t i = 0; i < n; ++i) s += x[i] * x[i];
  return sqrt(s);
}
double or_norm_l8(const double* x, int64_t n) {
  double s = 0.;
#pragma omp parallel for schedule(static) reduction(max : s) if (g_parvec)
  for (int64_t i = 0; i < n; ++i) s = fmax(s, fabs(x[i]));
  return s;
}
double or_dot(const double* x, const double* y, int64_t n) {
  double d = 0.;
#pragma omp parallel for schedule(static) reduction(+ : d) if (g_parvec)
  for (int64_t i = 0; i < n; ++i) d += x[i] * y[i];
  return d;
}
void or_equ(double* dst, double s, const double* src, int64_t n) {
#pragma omp parallel for schedule(static) if (g_parvec)
  for (int64_t i = 0; i < n; ++i) dst[i] = s * src[i];
}
void or_add(double* dst, double s, const double* src, int64_t n) {
#pragma omp parallel for schedule(static) if (g_parvec)
  for (int64_t i = 0; i < n; ++i) dst[i] += s * src[i];
}
void or_sadd(double* y, double a, double d, const double* v, int64_t n) {
#pragma omp parallel for schedule(static) if (g_parvec)
  for (int64_t i = 0; i < n; ++i) y[i] = a * y[i] + d * v[i];
}
void or_zero(double* x, int64_t n) { memset(x, 0, sizeof(double) * (size_t)n); }

/* ---------------------------------------------------------------- transfer */

typedef struct {
  int32_t r, c;
} or_pair;

static int pair_cmp(const void* a, const void* b) {
  const or_pair* x = (const or_pair*)a;
  const or_pair* y = (const or_pair*)b;
  if (x->r != y->r) return x->r < y->r ? -1 : 1;
  if (x->c != y->c) return x->c < y->c ? -1 : 1;
  return 0;
}

/* sorted, de-duplicated CSR from pairs: SparseStructure rows are std::set<int>
 * (src/LinAlg/sparsestructure.h:34-129), copied in order by ColumnStencil. */
static void pairs_to_csr(or_pair* P, int64_t np, int nrows, int32_t** rowp, int32_t** colp,
                         int64_t* nnz) {
  qsort(P, (size_t)np, sizeof(or_pair), pair_cmp);
  int64_t m = 0;
  for (int64_t k = 0; k < np; ++k)
    if (m == 0 || P[k].r != P[m - 1].r || P[k].c != P[m - 1].c) P[m++] = P[k];
  int32_t* row = (int32_t*)calloc((size_t)nrows + 1, sizeof(int32_t));
  int32_t* col = (int32_t*)malloc(sizeof(int32_t) * (size_t)(m > 0 ? m : 1));
  for (int64_t k = 0; k < m; ++k) {
    row[P[k].r + 1]++;
    col[k] = P[k].c;
  }
  for (int i = 0; i < nrows; ++i) row[i + 1] += row[i];
  *rowp = row;
  *colp = col;
  *nnz = m;
}

/* MgInterpolatorMatrix::BasicInit (src/Discretization/Q1/mginterpolatormatrix.cc:34-121) */
or_transfer* or_transfer_create(int ncoarse, const int32_t* c2f, int n2,
                                const int32_t* z, int n4, const int32_t* v, int n8,
                                const int32_t* a) {
  int nfine = -1;
  for (int i = 0; i < ncoarse; ++i) nfine = c2f[i] > nfine ? c2f[i] : nfine;
  for (int i = 0; i < n2; ++i) nfine = z[3 * i] > nfine ? z[3 * i] : nfine;
  for (int i = 0; i < n4; ++i) nfine = v[5 * i] > nfine ? v[5 * i] : nfine;
  for (int i = 0; i < n8; ++i) nfine = a[9 * i] > nfine ? a[9 * i] : nfine;
  nfine++;
  int64_t np = (int64_t)ncoarse + 2 * (int64_t)n2 + 4 * (int64_t)n4 + 8 * (int64_t)n8;
  or_pair* F = (or_pair*)malloc(sizeof(or_pair) * (size_t)(np > 0 ? np : 1));
  or_pair* C = (or_pair*)malloc(sizeof(or_pair) * (size_t)(np > 0 ? np : 1));
  double* wf = (double*)malloc(sizeof(double) * (size_t)nfine);
  for (int i = 0; i < nfine; ++i) wf[i] = 1.0; /* in I2C list */
  int64_t k = 0;
  for (int i = 0; i < ncoarse; ++i, ++k) {
    F[k].r = c2f[i]; F[k].c = i;
    C[k].r = i; C[k].c = c2f[i];
  }
  for (int i = 0; i < n2; ++i)
    for (int j = 0; j < 2; ++j, ++k) {
      F[k].r = z[3 * i]; F[k].c = z[3 * i + 1 + j];
      C[k].r = z[3 * i + 1 + j]; C[k].c = z[3 * i];
    }
  for (int i = 0; i < n4; ++i)
    for (int j = 0; j < 4; ++j, ++k) {
      F[k].r = v[5 * i]; F[k].c = v[5 * i + 1 + j];
      C[k].r = v[5 * i + 1 + j]; C[k].c = v[5 * i];
    }
  for (int i = 0; i < n8; ++i)
    for (int j = 0; j < 8; ++j, ++k) {
      F[k].r = a[9 * i]; F[k].c = a[9 * i + 1 + j];
      C[k].r = a[9 * i + 1 + j]; C[k].c = a[9 * i];
    }
  /* weight lookup order: I2, then I4, then I8, else 1 */
  for (int i = n8 - 1; i >= 0; --i) wf[a[9 * i]] = 0.125;
  for (int i = n4 - 1; i >= 0; --i) wf[v[5 * i]] = 0.25;
  for (int i = n2 - 1; i >= 0; --i) wf[z[3 * i]] = 0.5;

  or_transfer* T = (or_transfer*)calloc(1, sizeof(or_transfer));
  T->ncoarse = ncoarse;
  T->nfine = nfine;
  int64_t nf, nc;
  pairs_to_csr(F, np, nfine, &T->frow, &T->fcol, &nf);
  pairs_to_csr(C, np, ncoarse, &T->crow, &T->ccol, &nc);
  T->fval = (double*)malloc(sizeof(double) * (size_t)(nf > 0 ? nf : 1));
  T->cval = (double*)malloc(sizeof(double) * (size_t)(nc > 0 ? nc : 1));
  for (int r = 0; r < nfine; ++r)
    for (int pos = T->frow[r]; pos < T->frow[r + 1]; ++pos) T->fval[pos] = wf[r];
  for (int r = 0; r < ncoarse; ++r)
    for (int pos = T->crow[r]; pos < T->crow[r + 1]; ++pos) T->cval[pos] = wf[T->ccol[pos]];
  free(F);
  free(C);
  free(wf);
  return T;
}

void or_transfer_free(or_transfer* T) {
  if (!T) return;
  free(T->crow); free(T->ccol); free(T->cval);
  free(T->frow); free(T->fcol); free(T->fval);
  free(T);
}

int64_t or_transfer_nnz(const or_transfer* T) { return T->crow[T->ncoarse]; }

void or_transfer_export(const or_transfer* T, int32_t* crow, int32_t* ccol, double* cval,
                        int32_t* frow, int32_t* fcol, double* fval) {
  int64_t nnz = or_transfer_nnz(T);
  memcpy(crow, T->crow, sizeof(int32_t) * (size_t)(T->ncoarse + 1));
  memcpy(ccol, T->ccol, sizeof(int32_t) * (size_t)nnz);
  memcpy(cval, T->cval, sizeof(double) * (size_t)nnz);
  memcpy(frow, T->frow, sizeof(int32_t) * (size_t)(T->nfine + 1));
  memcpy(fcol, T->fcol, sizeof(int32_t) * (size_t)nnz);
  memcpy(fval, T->fval, sizeof(double) * (size_t)nnz);
}

/* MgInterpolatorMatrix::restrict_zero (mginterpolatormatrix.cc:123-133) */
void or_restrict_zero(const or_transfer* T, int ncomp, double* uc, const double* uf) {
#pragma omp parallel for
  for (int i = 0; i < T->ncoarse; ++i) {
    double* p = uc + (size_t)i * ncomp;
    for (int c = 0; c < ncomp; ++c) p[c] = 0.;
    for (int pos = T->crow[i]; pos < T->crow[i + 1]; ++pos) {
      const double* q = uf + (size_t)T->ccol[pos] * ncomp;
      for (int c = 0; c < ncomp; ++c) p[c] += T->cval[pos] * q[c];
    }
  }
}

/* MgInterpolatorMatrix::prolongate_add (mginterpolatormatrix.cc:137-147) */
void or_prolongate_add(const or_transfer* T, int ncomp, double* uf, const double* uc) {
#pragma omp parallel for
  for (int i = 0; i < T->nfine; ++i) {
    double* p = uf + (size_t)i * ncomp;
    for (int pos = T->frow[i]; pos < T->frow[i + 1]; ++pos) {
      const double* q = uc + (size_t)T->fcol[pos] * ncomp;
      for (int c = 0; c < ncomp; ++c) p[c] += T->fval[pos] * q[c];
    }
  }
}

/* ----------------------------------------------------------- hanging nodes */
/* Q1: HNStructureQ12d (src/Discretization/Q1/hnstructureq12d.cc:196-245),
 *     HNStructureQ13d (src/Discretization/Q1/hnstructureq13d.cc:145-185);
 * Q2: HNStructureQ22d (src/Discretization/Q2/hnstructureq22d.cc:58-80),
 *     HNStructureQ23d (src/Discretization/Q2/hnstructureq23d.cc:33-47,120-156). */
static const double Q1W[3] = {0.5, 0.5, 0.};
static const double Q2W[3] = {0.375, 0.75, -0.125};

void or_hn_average(const or_hn* H, int N, double* u) {
  const int q2 = (H->kind == 22 || H->kind == 23);
  if (H->kind == 23) {
    double fw[9];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) fw[3 * i + j] = Q2W[i] * Q2W[j];
    for (int k = 0; k < H->nf; ++k) {
      const int32_t* f = H->faces + 10 * k;
      double* p = u + (size_t)f[0] * N;
      const int32_t* g = f + 1;
      for (int c = 0; c < N; ++c)
        p[c] = fw[0] * u[(size_t)g[0] * N + c] + fw[1] * u[(size_t)g[1] * N + c] +
               fw[2] * u[(size_t)g[2] * N + c];
      for (int c = 0; c < N; ++c)
        p[c] += fw[3] * u[(size_t)g[3] * N + c] + fw[4] * u[(size_t)g[4] * N + c] +
                fw[5] * u[(size_t)g[5] * N + c];
      for (int c = 0; c < N; ++c)
        p[c] += fw[6] * u[(size_t)g[6] * N + c] + fw[7] * u[(size_t)g[7] * N + c] +
                fw[8] * u[(size_t)g[8] * N + c];
    }
  }
  for (int k = 0; k < H->ne; ++k) {
    const int32_t* e = H->edges + 4 * k;
    double* p = u + (size_t)e[0] * N;
    if (q2) {
      for (int c = 0; c < N; ++c)
        p[c] = Q2W[0] * u[(size_t)e[1] * N + c] + Q2W[1] * u[(size_t)e[2] * N + c] +
               Q2W[2] * u[(size_t)e[3] * N + c];
    } else {
      for (int c = 0; c < N; ++c)
        p[c] = Q1W[0] * u[(size_t)e[1] * N + c] + Q1W[1] * u[(size_t)e[2] * N + c];
    }
  }
  if (H->kind == 13) {
    for (int k = 0; k < H->nf; ++k) {
      const int32_t* f = H->faces + 10 * k;
      const int32_t* g = f + 1;
      double* p = u + (size_t)f[0] * N;
      for (int c = 0; c < N; ++c)
        p[c] = 0.25 * u[(size_t)g[0] * N + c] + 0.25 * u[(size_t)g[1] * N + c] +
               0.25 * u[(size_t)g[3] * N + c] + 0.25 * u[(size_t)g[4] * N + c];
    }
  }
}

static inline void add_node(double* u, int N, int dst, double w, int src) {
  for (int c = 0; c < N; ++c) u[(size_t)dst * N + c] += w * u[(size_t)src * N + c];
}
static inline void zero_node(double* u, int N, int i) {
  for (int c = 0; c < N; ++c) u[(size_t)i * N + c] = 0.;
}

void or_hn_distribute(const or_hn* H, int N, double* u) {
  const int q2 = (H->kind == 22 || H->kind == 23);
  if (H->kind == 23) {
    double fw[9];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) fw[3 * i + j] = Q2W[i] * Q2W[j];
    for (int k = 0; k < H->nf; ++k) {
      const int32_t* f = H->faces + 10 * k;
      for (int j = 0; j < 9; ++j) add_node(u, N, f[1 + j], fw[j], f[0]);
      zero_node(u, N, f[0]);
    }
  }
  for (int k = 0; k < H->ne; ++k) {
    const int32_t* e = H->edges + 4 * k;
    if (q2) {
      for (int j = 0; j < 3; ++j) add_node(u, N, e[1 + j], Q2W[j], e[0]);
    } else {
      add_node(u, N, e[1], Q1W[0], e[0]);
      add_node(u, N, e[2], Q1W[1], e[0]);
    }
    zero_node(u, N, e[0]);
  }
  if (H->kind == 13) {
    for (int k = 0; k < H->nf; ++k) {
      const int32_t* f = H->faces + 10 * k;
      add_node(u, N, f[1], 0.25, f[0]);
      add_node(u, N, f[2], 0.25, f[0]);
      add_node(u, N, f[4], 0.25, f[0]);
      add_node(u, N, f[5], 0.25, f[0]);
      zero_node(u, N, f[0]);
    }
  }
}

void or_hn_zero(const or_hn* H, int N, double* u) {
  for (int k = 0; k < H->ne; ++k) zero_node(u, N, H->edges[4 * k]);
  if (H->kind == 13 || H->kind == 23)
    for (int k = 0; k < H->nf; ++k) zero_node(u, N, H->faces[10 * k]);
}

/* CGDisc::StrongDirichletVectorZero (src/DofHandler/cgdisc.h:1162-1179) */
void or_dirichlet_zero(int nidx, const int32_t* idx, int ncomp, int ncomps,
                       const int32_t* comps, double* u) {
  for (int i = 0; i < nidx; ++i)
    for (int k = 0; k < ncomps; ++k) u[(size_t)idx[i] * ncomp + comps[k]] = 0.;
}

/* ------------------------------------------------------------------ CGInfo */
/* CGInfo::CGInfo / reset / check (src/Common/cginfo.cc:148-256) */
void or_cginfo_init(or_cginfo* I, double tol, double globaltol, int maxiter) {
  memset(I, 0, sizeof(*I));
  I->tol = tol;
  I->globaltol = globaltol;
  I->maxiter = maxiter;
  I->miniter = 0;
  I->breaktol = 1.e15;
  I->lastrate = 1.;
  or_cginfo_reset(I);
}

void or_cginfo_reset(or_cginfo* I) {
  I->residual = 0.;
  I->correction = 0.;
  I->iteration = 0;
  I->status = OR_WAITING;
  I->previousresidual = 1.;
  I->previouscorrection = 1.;
  I->rate = 0.;
  I->lastrate = 0.;
}

int or_cginfo_check(or_cginfo* I, double resi, double cori) {
  double res = fabs(resi), cor = fabs(cori);
  I->residual = res;
  I->correction = cor;
  if (I->status != OR_RUNNING) {
    I->firstresidual = res;
    I->previousresidual = res;
    I->previouscorrection = cor;
    I->aimedresidual = res * I->tol;
    I->status = OR_RUNNING;
  } else {
    I->totaliter++;
    I->lastrate = res / I->previousresidual;
    {
      double b = I->residual / I->firstresidual;
      double p = 1. / (double)(I->iteration > 1 ? I->iteration : 1);
      I->rate = pow(b, p);
    }
    I->previousresidual = res;
  }
  if (I->residual < fmax(I->aimedresidual, I->globaltol))
    I->status = OR_CONVERGED;
  else if (I->iteration >= I->maxiter)
    I->status = OR_TOO_MANY;
  else if (I->residual > I->breaktol)
    I->status = OR_EXPLODED;
  else if (I->residual > I->previousresidual)
    I->status = OR_DIVERGING;
  I->iteration++;
  if (I->status == OR_RUNNING) return 0;
  if (I->iteration < I->miniter) return 0;
  return 1;
}

/* ------------------------------------------------------------------ NLInfo */
/* NLInfo (src/Common/nlinfo.cc:100-356) */
void or_nlinfo_init(or_nlinfo* I, double tol, double globaltol, int maxiter) {
  memset(I, 0, sizeof(*I));
  I->tol = tol;
  I->globaltol = globaltol;
  I->maxiter = maxiter;
  I->breaktol = 1.e15;
  I->maxrelax = 11;
  I->rho = 0.3;
  I->linrho = 1.;
  I->maxresincrease = 1.e3;
  or_nlinfo_reset(I);
}

void or_nlinfo_reset(or_nlinfo* I) {
  I->rate = 0.;
  I->residual = 0.;
  I->correction = 0.;
  I->iteration = 0;
  I->status = OR_NL_WAITING;
  I->previousresidual = 1.;
  I->previouscorrection = 1.;
  I->relax = 0;
  I->laststepbad = 0;
}

int or_nlinfo_check_damping(or_nlinfo* I, int dampit, double res) {
  I->residual = fabs(res);
  I->relax = dampit;
  if (I->residual > 1.e20) return 2;
  if (!(I->residual <= 1.e20)) return 2;
  if (I->residual <= I->previousresidual) return 0;
  return 1;
}

int or_nlinfo_check(or_nlinfo* I, int iter, double resi, double cori) {
  double res = fabs(resi), cor = fabs(cori);
  int newiteration = 0;
  if (I->status == OR_NL_DIVERGED) return 1;
  if (I->status == OR_NL_STAGNATION) return 1;
  if (I->status == OR_NL_WAITING) {
    newiteration = 1;
  } else if (iter > I->iteration) {
    newiteration = 1;
    I->iteration++;
  }
  I->status = OR_NL_RUNNING;
  int thisstepbad = 0;
  if (!I->iteration) {
    I->residual = res;
    I->previousresidual = res;
    I->correction = cor;
    I->firstresidual = res;
    I->aimedresidual = res * I->tol;
  } else {
    double r0 = I->residual, c0 = I->correction;
    I->residual = res;
    I->correction = cor;
    {
      double b = I->residual / I->firstresidual;
      double p = 1. / (double)(I->iteration > 1 ? I->iteration : 1);
      I->rate = pow(b, p);
      if (I->iteration > 1)
        I->lastrate = I->residual / I->previousresidual;
      else
        I->lastrate = I->rate;
    }
    if (I->residual > 2. * I->previousresidual) thisstepbad = 1;
    if (newiteration) {
      I->previousresidual = r0;
      I->previouscorrection = c0;
    }
  }
  if (thisstepbad && I->laststepbad)
    I->status = OR_NL_DIVERGED;
  else if (I->residual < fmax(I->aimedresidual, I->globaltol))
    I->status = OR_NL_CONVERGED;
  else if (I->iteration >= I->maxiter && I->rate < 1.)
    I->status = OR_NL_SLOW;
  else if (I->iteration >= I->maxiter)
    I->status = OR_NL_STAGNATION;
  else if (I->residual > I->breaktol)
    I->status = OR_NL_EXPLODED;
  else if (I->iteration && I->lastrate > I->maxresincrease)
    I->status = OR_NL_EXPLODED;
  I->laststepbad = thisstepbad;
  if (I->status == OR_NL_RUNNING) return 0;
  if (I->iteration < I->miniter) return 0;
  I->totaliter += I->iteration;
  return 1;
}

/* ---------------------------------------------------------------- direct LU */
/* Dense LU with partial pivoting for direct coarse levels.  The reference
 * uses UMFPACK here (UmfIlu::Solve, src/LinAlg/umfilu.cc), which is absent:
 * this is a mathematically exact stand-in, parity with UMFPACK unpinned. */
static int dense_lu(int n, double* a, int32_t* piv) {
  for (int k = 0; k < n; ++k) {
    int p = k;
    double m = fabs(a[(size_t)k * n + k]);
    for (int i = k + 1; i < n; ++i)
      if (fabs(a[(size_t)i * n + k]) > m) {
        m = fabs(a[(size_t)i * n + k]);
        p = i;
      }
    piv[k] = p;
    if (m == 0.) return -1;
    if (p != k)
      for (int j = 0; j < n; ++j) {
        double t = a[(size_t)k * n + j];
        a[(size_t)k * n + j] = a[(size_t)p * n + j];
        a[(size_t)p * n + j] = t;
      }
    for (int i = k + 1; i < n; ++i) {
      double l = a[(size_t)i * n + k] / a[(size_t)k * n + k];
      a[(size_t)i * n + k] = l;
      for (int j = k + 1; j < n; ++j) a[(size_t)i * n + j] -= l * a[(size_t)k * n + j];
    }
  }
  return 0;
}

static void dense_lu_solve(int n, const double* a, const int32_t* piv, double* x) {
  for (int k = 0; k < n; ++k)
    if (piv[k] != k) {
      double t = x[k];
      x[k] = x[piv[k]];
      x[piv[k]] = t;
    }
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < i; ++j) x[i] -= a[(size_t)i * n + j] * x[j];
  for (int i = n - 1; i >= 0; --i) {
    for (int j = i + 1; j < n; ++j) x[i] -= a[(size_t)i * n + j] * x[j];
    x[i] /= a[(size_t)i * n + i];
  }
}

/* --------------------------------------------------------------- multigrid */

or_mg* or_mg_create(int nlevels) {
  or_mg* M = (or_mg*)calloc(1, sizeof(or_mg));
  M->nlevels = nlevels;
  M->L = (or_level*)calloc((size_t)nlevels, sizeof(or_level));
  M->mgtype = 'V';
  M->iterpre = 4;
  M->iterpost = 4;
  M->iterexact = 10;
  M->omega = 1.;
  M->mgomega = 1.;
  M->smoother = 0;
  M->prec_maxiter = 1;
  M->prec_tol = 1.e-12;
  M->prec_globaltol = 1.e-12;
  return M;
}

void or_mg_free(or_mg* M) {
  if (!M) return;
  for (int l = 0; l < M->nlevels; ++l) {
    if (M->u) free(M->u[l]);
    if (M->b) free(M->b[l]);
    if (M->v) free(M->v[l]);
    free(M->L[l].lu);
    free(M->L[l].piv);
  }
  free(M->u);
  free(M->b);
  free(M->v);
  free(M->L);
  free(M);
}

static int64_t lsize(const or_mg* M, int l) {
  return (int64_t)M->L[l].A.n * M->L[l].A.ncomp;
}

int or_mg_setup(or_mg* M) {
  M->u = (double**)calloc((size_t)M->nlevels, sizeof(double*));
  M->b = (double**)calloc((size_t)M->nlevels, sizeof(double*));
  M->v = (double**)calloc((size_t)M->nlevels, sizeof(double*));
  for (int l = 0; l < M->nlevels; ++l) {
    int64_t n = lsize(M, l);
    M->u[l] = (double*)calloc((size_t)n, sizeof(double));
    M->b[l] = (double*)calloc((size_t)n, sizeof(double));
    M->v[l] = (double*)calloc((size_t)n, sizeof(double));
    or_level* L = &M->L[l];
    if (L->direct) {
      int N = (int)n, B = L->A.ncomp;
      L->lu = (double*)calloc((size_t)N * N, sizeof(double));
      L->piv = (int32_t*)malloc(sizeof(int32_t) * (size_t)N);
      for (int i = 0; i < L->A.n; ++i)
        for (int pos = L->A.rowptr[i]; pos < L->A.rowptr[i + 1]; ++pos)
          for (int r = 0; r < B; ++r)
            for (int c = 0; c < B; ++c)
              L->lu[(size_t)(i * B + r) * N + (L->A.col[pos] * B + c)] +=
                  L->A.val[(size_t)pos * B * B + r * B + c];
      if (dense_lu(N, L->lu, L->piv)) return -1;
    }
  }
  return 0;
}

/* StdSolver::smooth (src/Solver/stdsolver.cc:951-986) */
static void smooth(or_mg* M, int l, int niter, double* x, const double* y, double* h) {
  const or_matrix* A = &M->L[l].A;
  const int64_t n = lsize(M, l);
  for (int it = 0; it < niter; ++it) {
    if (M->smoother == 0) {
      or_residual(A, h, x, y);
      or_jacobi(A, h);
      or_add(x, M->omega, h, n);
    } else if (M->smoother == 1) {
      or_residual(A, h, x, y);
      or_add(x, M->omega, h, n);
    } else if (M->smoother == 3) {
      or_residual(A, h, x, y);
      or_vanka_solve((const or_vanka*)M->L[l].vanka, h);
      or_add(x, M->omega, h, n);
    } else if (M->smoother == 4) {
      or_residual(A, h, x, y);
      or_ilu_solve((const or_ilu*)M->L[l].ilu, h);
      or_add(x, M->omega, h, n);
    }
  }
}

/* StdSolver::smooth_exact (stdsolver.cc:1003-1025) */
static void smooth_exact(or_mg* M, int l, double* x, const double* y, double* h) {
  or_level* L = &M->L[l];
  if (L->direct) {
    int64_t n = lsize(M, l);
    memcpy(x, y, sizeof(double) * (size_t)n);
    dense_lu_solve((int)n, L->lu, L->piv, x);
  } else {
    smooth(M, l, M->iterexact, x, y, h);
  }
}

static void hn_distribute(or_mg* M, int l, double* x) {
  or_level* L = &M->L[l];
  if (L->hn.ne || L->hn.nf) or_hn_distribute(&L->hn, L->A.ncomp, x);
}
static void hn_average(or_mg* M, int l, double* x) {
  or_level* L = &M->L[l];
  if (L->hn.ne || L->hn.nf) or_hn_average(&L->hn, L->A.ncomp, x);
}
static void hn_zero(or_mg* M, int l, double* x) {
  or_level* L = &M->L[l];
  if (L->hn.ne || L->hn.nf) or_hn_zero(&L->hn, L->A.ncomp, x);
}
static void bnd_zero(or_mg* M, int l, double* x) {
  or_level* L = &M->L[l];
  or_dirichlet_zero(L->ndir, L->dir, L->A.ncomp, L->ndircomp, L->dircomp, x);
}

/* StdMultiLevelSolver::mgstep (src/Solver/stdmultilevelsolver.cc:462-531).
 * The vectors of level l are (u_l, b_l, v_l); on the finest level they are the
 * caller's u and the _mg0_/_mg1_ vectors. */
static void mgstep(or_mg* M, double* res, double* rw, int l, int finelevel, int coarselevel,
                   char* p0, char p, double** U, double** Bv, double** Vv) {
  M->mgsteps++;
  double *u = U[l], *b = Bv[l], *v = Vv[l];
  const or_matrix* A = &M->L[l].A;
  if (l == coarselevel) {
    if (p == 'F') *p0 = 'V';
    smooth_exact(M, l, u, b, v);
    if (l == finelevel) {
      or_residual(A, v, u, b);
      res[l] = or_norm(v, lsize(M, l));
    }
  } else {
    smooth(M, l, M->iterpre, u, b, v);
    or_residual(A, v, u, b);
    or_restrict_zero(M->L[l].T, A->ncomp, Bv[l - 1], v);
    if (!(M->plant & OR_PLANT_NO_HN_DISTRIBUTE)) hn_distribute(M, l - 1, Bv[l - 1]);
    bnd_zero(M, l - 1, Bv[l - 1]);
    or_zero(U[l - 1], lsize(M, l - 1));
    int j = 0;
    if (*p0 == 'V') j = 1;
    if (*p0 == 'W') j = 2;
    if (*p0 == 'F') j = 3;
    for (int i = 0; i < j; ++i)
      mgstep(M, res, rw, l - 1, finelevel, coarselevel, p0, p, U, Bv, Vv);
    if (l == 0 && p == 'F') *p0 = 'W';
    rw[l] = or_norm(U[l - 1], lsize(M, l - 1));
    or_zero(v, lsize(M, l));
    if (!(M->plant & OR_PLANT_NO_HN_AVERAGE)) hn_average(M, l - 1, U[l - 1]);
    or_prolongate_add(M->L[l].T, A->ncomp, v, U[l - 1]);
    hn_zero(M, l - 1, U[l - 1]);
    if (!(M->plant & OR_PLANT_NO_HN_ZERO_FINE)) hn_zero(M, l, v);
    bnd_zero(M, l, v);
    or_add(u, M->mgomega, v, lsize(M, l));
    smooth(M, l, M->iterpost, u, b, v);
    or_residual(A, v, u, b);
    res[l] = or_norm(v, lsize(M, l));
  }
}

/* StdMultiLevelSolver::LinearMg (src/Solver/stdmultilevelsolver.cc:411-458) */
void or_linear_mg(or_mg* M, int finelevel, int coarselevel, double* u, const double* f,
                  or_cginfo* info) {
  int clevel = coarselevel;
  for (int level = coarselevel; level < M->nlevels; ++level)
    if (M->L[level].direct) clevel = level;
  if (clevel > finelevel) clevel = finelevel;
  double* res = (double*)calloc((size_t)M->nlevels, sizeof(double));
  double* rw = (double*)calloc((size_t)M->nlevels, sizeof(double));
  double** U = (double**)malloc(sizeof(double*) * (size_t)M->nlevels);
  for (int l = 0; l < M->nlevels; ++l) U[l] = M->u[l];
  U[finelevel] = u;
  double* mg0 = M->b[finelevel];
  double* mg1 = M->v[finelevel];
  const or_matrix* A = &M->L[finelevel].A;
  const int64_t n = lsize(M, finelevel);
  or_equ(mg0, 1., f, n);
  or_residual(A, mg1, u, mg0);
  res[finelevel] = or_norm(mg1, n);
  rw[finelevel] = 0.;
  or_cginfo_check(info, res[finelevel], rw[finelevel]);
  int reached = 0;
  while (!reached) {
    char p = M->mgtype;
    char p0 = p;
    if (p == 'F') p0 = 'W';
    mgstep(M, res, rw, finelevel, finelevel, clevel, &p0, p, U, M->b, M->v);
    reached = or_cginfo_check(info, res[finelevel], rw[finelevel]);
  }
  free(res);
  free(rw);
  free(U);
}

/* StdMultiLevelSolver::precondition (stdmultilevelsolver.cc:940-952) */
void or_precondition(or_mg* M, int coarselevel, double* x, const double* y) {
  or_cginfo prec;
  or_cginfo_init(&prec, M->prec_tol, M->prec_globaltol, M->prec_maxiter);
  or_cginfo_reset(&prec);
  or_cginfo_check(&prec, 0., 0.);
  int clevel = coarselevel < 0 ? M->nlevels - 1 : coarselevel;
  or_linear_mg(M, M->nlevels - 1, clevel, x, y, &prec);
}

/* ------------------------------------------------------------------- GMRES */
/* GMRES::solve / givens_rotation / orthogonalization / reortho_test /
 * solution (src/LinAlg/gmres.cc:85-260), left preconditioning. */
int or_gmres_solve(or_mg* M, int coarselevel, int vmax, double* x, const double* b,
                   or_cginfo* info) {
  const int fl = M->nlevels - 1;
  const or_matrix* A = &M->L[fl].A;
  const int64_t n = lsize(M, fl);
  double** mem = (double**)calloc((size_t)vmax + 1, sizeof(double*));
  double** H = (double**)calloc((size_t)vmax, sizeof(double*));
  double* gamma = (double*)calloc((size_t)vmax + 1, sizeof(double));
  double* ci = (double*)calloc((size_t)vmax, sizeof(double));
  double* si = (double*)calloc((size_t)vmax, sizeof(double));
  for (int i = 0; i < vmax; ++i) H[i] = (double*)calloc((size_t)i + 2, sizeof(double));
  for (int i = 0; i < 3 && i <= vmax; ++i) mem[i] = (double*)calloc((size_t)n, sizeof(double));
  double* p = (double*)calloc((size_t)n, sizeof(double));
  int k0 = info->iteration;
  int reached = 0;
  double* v = mem[0];
  or_equ(p, 1., v, n);
  /* residualgmres (stdsolver.cc:743-754): p = A x; p = -p + b; Dirichlet 0 */
  or_zero(p, n);
  or_vmult(A, p, x, 1.);
  or_sadd(p, -1., 1., b, n);
  bnd_zero(M, fl, p);
  or_precondition(M, coarselevel, v, p);
  double rho = sqrt(or_dot(v, v, n));
  int dim = 0;
  if (rho == 0.) {
    reached = 1;
    goto done;
  }
  gamma[0] = rho;
  or_equ(v, 1. / rho, v, n);
  while (dim < vmax - 1 && !reached) {
    dim++;
    int last = dim - 1;
    if (!mem[dim]) mem[dim] = (double*)calloc((size_t)n, sizeof(double));
    double* vv = mem[dim];
    or_zero(p, n);
    or_vmult(A, p, mem[last], 1.);
    or_precondition(M, coarselevel, vv, p);
    double* h = H[last];
    for (int pass = 0; pass < 2; ++pass) {
      for (int i = 0; i < dim; ++i) {
        double d = or_dot(vv, mem[i], n);
        h[i] += d;
        or_add(vv, -d, mem[i], n);
      }
      h[dim] = sqrt(or_dot(vv, vv, n));
      if (pass == 0) {
        /* reortho_test: oldnorm == oldnorm + 1e-3*s */
        double oldnorm = sqrt(or_dot(vv, vv, n));
        double newnorm = oldnorm + 1.e-3 * h[dim];
        if (oldnorm != newnorm) break;
      }
    }
    double s = h[dim];
    or_equ(vv, 1. / s, vv, n);
    /* givens_rotation(H[last], last) */
    for (int i = 0; i < last; i++) {
      double dummy = h[i];
      h[i] = ci[i] * dummy + si[i] * h[i + 1];
      h[i + 1] = -si[i] * dummy + ci[i] * h[i + 1];
    }
    double r = 1. / sqrt(h[last] * h[last] + h[last + 1] * h[last + 1]);
    si[last] = h[last + 1] * r;
    ci[last] = h[last] * r;
    h[last] = ci[last] * h[last] + si[last] * h[last + 1];
    gamma[last + 1] = -si[last] * gamma[last];
    gamma[last] *= ci[last];
    rho = fabs(gamma[dim]);
    reached = or_cginfo_check(info, rho, (double)(k0 + last));
  }
  /* solution: backward (src/Common/mult.h:130-141) on H1(i,j) = H[j][i] */
  {
    double* hs = (double*)calloc((size_t)(dim > 0 ? dim : 1), sizeof(double));
    for (int i = dim - 1; i >= 0; i--) {
      double s = gamma[i];
      for (int j = i + 1; j < dim; j++) s -= hs[j] * H[j][i];
      hs[i] = s / H[i][i];
    }
    for (int i = 0; i < dim; i++) or_add(x, hs[i], mem[i], n);
    free(hs);
  }
done:
  for (int i = 0; i <= vmax; ++i) free(mem[i]);
  for (int i = 0; i < vmax; ++i) free(H[i]);
  free(mem);
  free(H);
  free(gamma);
  free(ci);
  free(si);
  free(p);
  return reached;
}

/* ------------------------------------------------------------------ Newton */
/* NewtonResidual (stdmultilevelsolver.cc:608-621) with the element Form
 * replaced by the assembled matrix (the same linear operator). */
static double newton_residual(or_mg* M, double* y, const double* x, const double* b,
                              int norm_kind) {
  const int fl = M->nlevels - 1;
  const int64_t n = lsize(M, fl);
  or_equ(y, 1., b, n);
  or_vmult(&M->L[fl].A, y, x, -1.);
  bnd_zero(M, fl, y);
  return norm_kind == 0 ? or_norm_l8(y, n) : or_norm(y, n);
}

/* StdMultiLevelSolver::Solve + newton + NewtonLinearSolve + NewtonUpdate
 * (stdmultilevelsolver.cc:548-570, 646-673, 693-758, 832-856). */
int or_newton_solve(or_mg* M, int coarselevel, int vmax, double* u, const double* f,
                    or_cginfo* linfo, or_nlinfo* nl, int norm_kind, or_solve_report* rep) {
  const int fl = M->nlevels - 1;
  const int64_t n = lsize(M, fl);
  double* r = (double*)calloc((size_t)n, sizeof(double));
  double* w = (double*)calloc((size_t)n, sizeof(double));
  int64_t mg0 = M->mgsteps;
  int gm = 0, it, ndamp = 0, nexpl = 0;
  hn_average(M, fl, u);
  or_nlinfo_reset(nl);
  or_cginfo_reset(linfo);
  double rr = newton_residual(M, r, u, f, norm_kind);
  if (rep) rep->first_residual = rr;
  int reached = or_nlinfo_check(nl, 0, rr, 0.);
  for (it = 1; !reached; it++) {
    or_zero(w, n);
    or_cginfo_reset(linfo);
    or_zero(w, n);
    int k_before = linfo->iteration;
    or_gmres_solve(M, coarselevel, vmax, w, r, linfo);
    gm += linfo->iteration - k_before;
    /* NewtonUpdate */
    int lex = linfo->status == OR_EXPLODED;
    double nn = norm_kind == 0 ? or_norm_l8(w, n) : or_norm(w, n);
    double nr = or_norm(r, n);
    if (nn > 1.e30 || !(nn >= 0.) || nr > 1.e30 || !(nr >= 0.)) lex = 1;
    double rw;
    if (lex) {
      nl->status = OR_NL_DIVERGED;
      rw = nn;
    } else {
      double omega = 0.7, relax = 1.;
      or_add(u, relax, w, n);
      rr = newton_residual(M, r, u, f, norm_kind);
      for (int iter = 0; iter < nl->maxrelax; iter++) {
        int msg = or_nlinfo_check_damping(nl, iter, rr);
        if (msg == 0) break;
        if (msg == 1) {
          ndamp++;
          or_add(u, relax * (omega - 1.), w, n);
          rr = newton_residual(M, r, u, f, norm_kind);
          relax *= omega;
          continue;
        }
        nexpl = 1;
        or_add(u, -relax, w, n);
        relax = 0.;
        nl->status = OR_NL_DIVERGED;
        break;
      }
      rw = nn;
    }
    reached = or_nlinfo_check(nl, it, rr, rw);
  }
  hn_zero(M, fl, u);
  if (rep) {
    rep->newton_iterations = nl->iteration;
    rep->gmres_iterations = gm;
    rep->mgsteps = M->mgsteps - mg0;
    rep->status = nl->status;
    rep->last_residual = rr;
    rep->damping_steps = ndamp;
    rep->damping_exploded = nexpl;
  }
  free(r);
  free(w);
  return nl->status == OR_NL_CONVERGED ? 0 : 1;
}
