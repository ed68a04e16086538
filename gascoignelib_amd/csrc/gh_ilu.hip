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
// the levels of the upper part (level 0: rows without upper entries).  Each
// row's operations run in one thread in the reference's order, every product
// and difference rounded separately, so the factors and the FP64 solves equal
// the CPU ILU's bit for bit.
//
// The factorisation runs on the factors in CSR form (one thread walks its
// row).  The solves run on a sliced copy of each triangle: the rows of a level
// in slices of 64 (a level starts a new slice), entry j of the 64 rows of a
// slice stored side by side, the lower part in ascending and the upper part in
// descending column order (the order the solves take them), so a wavefront's
// loads of one step are contiguous.  The diagonal inverses are a separate
// array indexed by ILU row.  FP32 solves (the reduced-precision V-cycle) use
// FP32 copies of the sliced values.
#include <algorithm>
#include <functional>
#include <numeric>
#include <vector>

#include "gh_kernels.h"

using namespace gh;

namespace gh {
// one triangle of the factors in slices of 64 rows, levels aligned to slices
struct IluSlices {
  int64_t nslices = 0, nentries = 0;
  std::vector<int64_t> lslice;  // level L covers slices [lslice[L], lslice[L+1]) (host)
  int32_t* row = nullptr;       // nslices*64 ILU rows, -1 in padding
  int32_t* len = nullptr;       // nslices*64 entry counts
  int64_t* base = nullptr;      // nslices first entry
  int32_t* col = nullptr;       // nentries
  int32_t* map = nullptr;       // nentries: CSR block position, -1 in padding
  double* v64 = nullptr;        // nentries*B*B
  float* v32 = nullptr;
};
}  // namespace gh

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
  double* dinv64 = nullptr;   // n diagonal inverses (from val) by ILU row
  float* dinv32 = nullptr;
  gh::IluSlices lo, up;       // sliced lower (forward) and upper (backward) parts
  void* y = nullptr;          // work vector, n*B doubles (or floats)
  int32_t* rows_f = nullptr;  // rows grouped by forward level (factorisation)
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
    // a(i,j) -= a(i,k) a(k,j) for the upper entries j of row k that row i
    // holds (submult); both rows are sorted, so one merge finds them (the
    // entries of row i right of column k start at pk + 1)
    int32_t ph = pk + 1;
    for (int32_t pj = diag[k] + 1; pj < rowptr[k + 1]; ++pj) {
      const int32_t j = col[pj];
      while (ph < e && col[ph] < j) ++ph;
      if (ph == e) break;
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

// forward (:193-203) or backward (:207-222) for the slices [s0, s1) of one
// level: y_i -= a(i,c) y_c over the row's entries in the sliced order, then,
// backward, y_i = a(i,i)^-1 y_i (FMatrixBlock::vmult).  The permutations of
// solve (:226-236) are folded in: the forward pass starts row i from
// xin[perm[i]] (hin), the backward pass also writes its result to
// xout[perm[i]] (her), and the smoother's update x += omega h
// (StdSolver::smooth, src/Solver/stdsolver.cc:962-967) can follow there:
// mode 0 xout = h, 1 xout += omega h, 2 xout = omega h.
template <int B, typename T, typename VT, bool kUpper>
__global__ __launch_bounds__(kBlock) void k_ilu_sweep(int64_t s0, int64_t s1,
                                                      const int32_t* __restrict__ row,
                                                      const int32_t* __restrict__ len,
                                                      const int64_t* __restrict__ base,
                                                      const int32_t* __restrict__ col,
                                                      const VT* __restrict__ val,
                                                      const VT* __restrict__ dinv,
                                                      const int32_t* __restrict__ perm,
                                                      const T* xin, T* __restrict__ y, T* xout,
                                                      int mode, T omega) {
  constexpr int BB = B * B;
  // entries whose loads are issued before their updates (fewer for larger
  // blocks, whose values take B*B registers each)
  constexpr int kGroup = B == 1 ? 8 : (B == 2 ? 4 : 2);
  const int64_t s = s0 + (int64_t)blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
  if (s >= s1) return;
  const int lane = threadIdx.x & 63;
  const int32_t i = row[s * 64 + lane];
  if (i < 0) return;
  const int32_t cnt = len[s * 64 + lane];
  const int64_t b = base[s] + lane;
  T yi[B];
  if (kUpper) {
#pragma unroll
    for (int r = 0; r < B; ++r) yi[r] = y[(int64_t)i * B + r];
  } else {
    const int64_t src = (int64_t)perm[i] * B;
#pragma unroll
    for (int r = 0; r < B; ++r) yi[r] = xin[src + r];
  }
  int32_t j = 0;
  for (; j + kGroup <= cnt; j += kGroup) {
    T yc[kGroup][B];
    VT m[kGroup][BB];
#pragma unroll
    for (int g = 0; g < kGroup; ++g) {
      const int64_t e = b + (int64_t)(j + g) * 64;
      const int32_t c = col[e];
#pragma unroll
      for (int r = 0; r < B; ++r) yc[g][r] = y[(int64_t)c * B + r];
#pragma unroll
      for (int x = 0; x < BB; ++x) m[g][x] = val[e * BB + x];
    }
#pragma unroll
    for (int g = 0; g < kGroup; ++g) block_subtract<B>(yi, m[g], yc[g]);
  }
  for (; j < cnt; ++j) {
    const int64_t e = b + (int64_t)j * 64;
    const int32_t c = col[e];
    T yc[B];
#pragma unroll
    for (int r = 0; r < B; ++r) yc[r] = y[(int64_t)c * B + r];
    block_subtract<B>(yi, val + e * BB, yc);
  }
  if (kUpper) {
    const VT* D = dinv + (int64_t)i * BB;
    const int64_t dst = (int64_t)perm[i] * B;
#pragma unroll
    for (int r = 0; r < B; ++r) {
      T a = (T)0;
#pragma unroll
      for (int c = 0; c < B; ++c) a = add_rn(a, mul_rn((T)D[r * B + c], yi[c]));
      y[(int64_t)i * B + r] = a;
      if (mode == 0)
        xout[dst + r] = a;
      else if (mode == 1)
        xout[dst + r] = add_rn(xout[dst + r], mul_rn(omega, a));
      else
        xout[dst + r] = mul_rn(omega, a);
    }
  } else {
#pragma unroll
    for (int r = 0; r < B; ++r) y[(int64_t)i * B + r] = yi[r];
  }
}

// sliced values (and the diagonal inverses) from the CSR factors
__global__ __launch_bounds__(kBlock) void k_ilu_fill(int64_t m, int BB, const int32_t* __restrict__ map,
                                                     const double* __restrict__ val,
                                                     double* __restrict__ v64,
                                                     float* __restrict__ v32) {
  for (int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x; k < m; k += (int64_t)gridDim.x * kBlock) {
    const int32_t pos = map[k];
    for (int x = 0; x < BB; ++x) {
      const double v = pos >= 0 ? val[(int64_t)pos * BB + x] : 0.0;
      v64[k * BB + x] = v;
      v32[k * BB + x] = (float)v;
    }
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
int solve_levels(gh_ctx* ctx, const gh_ilu* I, bool fp64, const T* xin, T* y, T* xout, int mode,
                 T omega) {
  for (int part = 0; part < 2; ++part) {
    const IluSlices& S = part == 0 ? I->lo : I->up;
    const VT* val = fp64 ? (const VT*)S.v64 : (const VT*)S.v32;
    const VT* dinv = fp64 ? (const VT*)I->dinv64 : (const VT*)I->dinv32;
    for (size_t L = 0; L + 1 < S.lslice.size(); ++L) {
      const int64_t s0 = S.lslice[L], s1 = S.lslice[L + 1];
      const unsigned g = (unsigned)((s1 - s0 + kBlock / 64 - 1) / (kBlock / 64));
      if (part == 0)
        hipLaunchKernelGGL((k_ilu_sweep<B, T, VT, false>), dim3(g), dim3(kBlock), 0, ctx->stream,
                           s0, s1, S.row, S.len, S.base, S.col, val, dinv, I->p, xin, y, xout,
                           mode, omega);
      else
        hipLaunchKernelGGL((k_ilu_sweep<B, T, VT, true>), dim3(g), dim3(kBlock), 0, ctx->stream,
                           s0, s1, S.row, S.len, S.base, S.col, val, dinv, I->p, xin, y, xout,
                           mode, omega);
    }
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

// one triangle in slices: the rows of each level (in `rows`, levels at
// lptr) in slices of 64; entry j of a row is CSR position rowptr[i] + j
// (lower part, ascending) or rowptr[i+1] - 1 - j (upper part, descending)
int build_slices(gh_ctx* ctx, const std::vector<int32_t>& rows, const std::vector<int32_t>& lptr,
                 const std::vector<int32_t>& rowptr, const std::vector<int32_t>& col,
                 const std::vector<int32_t>& diag, bool upper, int BB, IluSlices* S) {
  const size_t nl = lptr.size() - 1;
  S->lslice.assign(nl + 1, 0);
  for (size_t L = 0; L < nl; ++L) S->lslice[L + 1] = S->lslice[L] + (lptr[L + 1] - lptr[L] + 63) / 64;
  const int64_t ns = S->lslice[nl];
  std::vector<int32_t> hrow(ns * 64, -1), hlen(ns * 64, 0);
  std::vector<int64_t> hbase(ns + 1, 0);
  auto count = [&](int32_t i) { return upper ? rowptr[i + 1] - 1 - diag[i] : diag[i] - rowptr[i]; };
  for (size_t L = 0; L < nl; ++L)
    for (int32_t k = lptr[L]; k < lptr[L + 1]; ++k) {
      const int64_t slot = S->lslice[L] * 64 + (k - lptr[L]);
      hrow[slot] = rows[k];
      hlen[slot] = count(rows[k]);
    }
  for (int64_t sl = 0; sl < ns; ++sl) {
    int32_t w = 0;
    for (int l = 0; l < 64; ++l) w = std::max(w, hlen[sl * 64 + l]);
    hbase[sl + 1] = hbase[sl] + (int64_t)w * 64;
  }
  const int64_t ne = hbase[ns];
  std::vector<int32_t> hcol(ne, 0), hmap(ne, -1);
  for (int64_t sl = 0; sl < ns; ++sl)
    for (int l = 0; l < 64; ++l) {
      const int32_t i = hrow[sl * 64 + l];
      for (int32_t j = 0; j < hlen[sl * 64 + l]; ++j) {
        const int32_t pos = upper ? rowptr[i + 1] - 1 - j : rowptr[i] + j;
        hcol[hbase[sl] + (int64_t)j * 64 + l] = col[pos];
        hmap[hbase[sl] + (int64_t)j * 64 + l] = pos;
      }
    }
  hbase.pop_back();
  S->nslices = ns;
  S->nentries = ne;
  GH_TRY(dupload(ctx, &S->row, hrow));
  GH_TRY(dupload(ctx, &S->len, hlen));
  GH_TRY(dupload(ctx, &S->base, hbase));
  GH_TRY(dupload(ctx, &S->col, hcol));
  GH_TRY(dupload(ctx, &S->map, hmap));
  if (hipMalloc((void**)&S->v64, sizeof(double) * std::max<int64_t>(ne * BB, 1)) != hipSuccess ||
      hipMalloc((void**)&S->v32, sizeof(float) * std::max<int64_t>(ne * BB, 1)) != hipSuccess)
    return fail(GH_ENOMEM, "gh_ilu_create: allocation failed");
  // the host vectors die with this call: the uploads must have landed
  GH_HIP(hipStreamSynchronize(ctx->stream));
  return GH_OK;
}

void free_slices(IluSlices* S) {
  for (void* ptr : {(void*)S->row, (void*)S->len, (void*)S->base, (void*)S->col, (void*)S->map,
                    (void*)S->v64, (void*)S->v32})
    if (ptr) (void)hipFree(ptr);
  *S = IluSlices();
}

void fill_launch(gh_ctx* ctx, int64_t m, int BB, const int32_t* map, const double* val, double* v64,
                 float* v32) {
  if (m <= 0) return;
  const unsigned g = (unsigned)std::min<int64_t>((m + kBlock - 1) / kBlock, 8192);
  hipLaunchKernelGGL(k_ilu_fill, dim3(g), dim3(kBlock), 0, ctx->stream, m, BB, map, val, v64, v32);
}

}  // namespace

namespace gh {
int ilu_apply(gh_ctx* ctx, const gh_ilu* I, int prec, const void* in, void* out, int mode,
              double omega);
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
  if (!rc) rc = build_slices(ctx, rows_f, I->lptr_f, rowptr, col, diag, false, BB, &I->lo);
  if (!rc) rc = build_slices(ctx, rows_b, I->lptr_b, rowptr, col, diag, true, BB, &I->up);
  if (!rc && (hipMalloc((void**)&I->dinv64, sizeof(double) * std::max<int64_t>(n * BB, 1)) != hipSuccess ||
              hipMalloc((void**)&I->dinv32, sizeof(float) * std::max<int64_t>(n * BB, 1)) != hipSuccess ||
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
  if (!rc) {
    fill_launch(ctx, I->lo.nentries, BB, I->lo.map, I->val, I->lo.v64, I->lo.v32);
    fill_launch(ctx, I->up.nentries, BB, I->up.map, I->val, I->up.v64, I->up.v32);
    fill_launch(ctx, n, BB, I->diag, I->val, I->dinv64, I->dinv32);
    if (hipGetLastError() != hipSuccess) rc = fail(GH_EHIP, "gh_ilu_create: fill");
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
                    (void*)I->val, (void*)I->dinv64, (void*)I->dinv32, I->y, (void*)I->rows_f,
                    (void*)I->d_flag})
    if (ptr) (void)hipFree(ptr);
  free_slices(&I->lo);
  free_slices(&I->up);
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
  return ilu_apply(ctx, I, x->prec, x->d, x->d, 0, 1.0);
}

}  // extern "C"

namespace gh {

// SparseBlockIlu::solve (:226-236) of `in`, written to `out` (mode 0; `in`
// and `out` may be one vector), or added to it scaled by omega (mode 1), or
// stored scaled (mode 2); FP64 or FP32 vectors
int ilu_apply(gh_ctx* ctx, const gh_ilu* I, int prec, const void* in, void* out, int mode,
              double omega) {
  if (I->n == 0) return GH_OK;
  if (prec == GH_F64) {
    const double* xi = (const double*)in;
    double *y = (double*)I->y, *xo = (double*)out;
    switch (I->B) {
      case 1: GH_TRY((solve_levels<1, double, double>(ctx, I, true, xi, y, xo, mode, omega))); break;
      case 2: GH_TRY((solve_levels<2, double, double>(ctx, I, true, xi, y, xo, mode, omega))); break;
      case 3: GH_TRY((solve_levels<3, double, double>(ctx, I, true, xi, y, xo, mode, omega))); break;
      default: GH_TRY((solve_levels<4, double, double>(ctx, I, true, xi, y, xo, mode, omega))); break;
    }
  } else {
    const float* xi = (const float*)in;
    float *y = (float*)I->y, *xo = (float*)out;
    const float om = (float)omega;
    switch (I->B) {
      case 1: GH_TRY((solve_levels<1, float, float>(ctx, I, false, xi, y, xo, mode, om))); break;
      case 2: GH_TRY((solve_levels<2, float, float>(ctx, I, false, xi, y, xo, mode, om))); break;
      case 3: GH_TRY((solve_levels<3, float, float>(ctx, I, false, xi, y, xo, mode, om))); break;
      default: GH_TRY((solve_levels<4, float, float>(ctx, I, false, xi, y, xo, mode, om))); break;
    }
  }
  return GH_OK;
}

}  // namespace gh
