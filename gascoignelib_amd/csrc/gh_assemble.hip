// Device assembly of Q1 operators and of the element residual (Form) for
// -div(k grad u) = f on axis-aligned box cells, one coefficient per cell.
//
// Reference: CGDisc::Matrix (src/DofHandler/cgdisc.h:599-635) adds every
// cell's element matrix (elementintegrator.xx:350-397) into the global matrix
// in ascending cell order, then SparseBlockMatrix::dirichlet
// (src/LinAlg/sparseblockmatrix.xx:205-218) makes the Dirichlet rows and
// columns identity; CudaSolver::AssembleMatrix (cuda/cudasolver.cc:114-128)
// then uploads the host matrix and inverts its diagonal on every assembly.
// Here the values are formed on the device, directly in the matrix's own
// numbering and layout: one thread per row sums, for each of the row's
// entries, the contributions of the row node's cells in ascending cell order,
// k_c * K^c_ab with K^c the unit-coefficient Q1 element matrix formed exactly
// (tensor products of the 1D stiffness s and mass m matrices:
// K = s m m + m s m + m m s in 3D, s m + m s in 2D, each product and sum
// rounded separately in that order), which is what the host producer
// (structure.cc, assemble) computes; the two agree bit for bit, and with the
// reference's Gauss-rule values to rounding.
//
// CGDisc::Form (cgdisc.h:415-448) is the residual counterpart: y += d * sum
// over cells of the element residual.  gh_form_q1 forms r = f - A(u) cell by
// cell from u without any matrix.  gh_assemble_q2 / gh_form_q2 do the same
// for the triquadratic patch elements of CGDiscQ23d.
#include <algorithm>
#include <cmath>
#include <vector>

#include "gh_kernels.h"

using namespace gh;

namespace {

// Corner a of a hex at (cx, cy, cz): 0 (0,0,0), 1 (1,0,0), 2 (1,1,0), 3 (0,1,0),
// then the same at z = 1 (Gascoigne's order); quads use corners 0..3.
__device__ __forceinline__ int cx(int a) { return (a & 1) ^ ((a >> 1) & 1); }
__device__ __forceinline__ int cy(int a) { return (a >> 1) & 1; }
__device__ __forceinline__ int cz(int a) { return (a >> 2) & 1; }

// 1D Q1 stiffness and mass on an interval of length h
__device__ __forceinline__ double s1(int p, int q, double h) { return (p == q ? 1.0 : -1.0) / h; }
__device__ __forceinline__ double m1(int p, int q, double h) {
  return mul_rn(h, p == q ? 2.0 : 1.0) / 6.0;
}

template <int DIM>
__device__ __forceinline__ double q1_entry(int a, int b, double hx, double hy, double hz) {
  if constexpr (DIM == 3) {
    const double t1 = mul_rn(mul_rn(s1(cx(a), cx(b), hx), m1(cy(a), cy(b), hy)), m1(cz(a), cz(b), hz));
    const double t2 = mul_rn(mul_rn(m1(cx(a), cx(b), hx), s1(cy(a), cy(b), hy)), m1(cz(a), cz(b), hz));
    const double t3 = mul_rn(mul_rn(m1(cx(a), cx(b), hx), m1(cy(a), cy(b), hy)), s1(cz(a), cz(b), hz));
    return add_rn(add_rn(t1, t2), t3);
  } else {
    (void)hz;
    return add_rn(mul_rn(s1(cx(a), cx(b), hx), m1(cy(a), cy(b), hy)),
                  mul_rn(m1(cx(a), cx(b), hx), s1(cy(a), cy(b), hy)));
  }
}

// One thread per row i of A (A's numbering; the cells' node of row i is
// perm[i], or i).  Entry (i, j): sum over the cells c of node perm[i], in
// ascending order, of k_c K^c_ab where a, b are the local corners of the two
// nodes in c.  Dirichlet rows are unit rows, Dirichlet columns zero.  Each
// (cell, corner) pair of the row must meet exactly one entry of the pattern,
// else flags[2] is set.  The row's cells (at most MAXC; rows of more cells
// take the generic loop) are read once into registers.
constexpr int kMaxRowCells = 8;

template <int DIM>
__global__ __launch_bounds__(kBlock) void k_assemble_q1(
    int64_t n, const int32_t* __restrict__ rowptr, const int32_t* __restrict__ col,
    const int32_t* __restrict__ perm, const int32_t* __restrict__ nptr,
    const int32_t* __restrict__ nlist, const int32_t* __restrict__ cells,
    const double* __restrict__ h, const double* __restrict__ kc,
    const uint8_t* __restrict__ isdir, double* __restrict__ val, int* flags) {
  constexpr int NPC = DIM == 3 ? 8 : 4;
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const int32_t ic = perm ? perm[i] : (int32_t)i;
  const int32_t c0 = nptr[ic], c1 = nptr[ic + 1];
  const int nc = c1 - c0;
  const bool drow = isdir && isdir[i];
  const int32_t q0 = rowptr[i], q1 = rowptr[i + 1];
  int matched = 0;
  if (nc <= kMaxRowCells) {
    int32_t node[kMaxRowCells][NPC];
    int loc[kMaxRowCells];
    double kk[kMaxRowCells], hx[kMaxRowCells], hy[kMaxRowCells], hz[kMaxRowCells];
#pragma unroll
    for (int e = 0; e < kMaxRowCells; ++e) {
      const int32_t ce = e < nc ? nlist[c0 + e] : 0;
      const int64_t c = ce >> 3;
      loc[e] = ce & 7;
      kk[e] = (e < nc && kc) ? kc[c] : 1.0;
      hx[e] = e < nc ? h[3 * c] : 1.0;
      hy[e] = e < nc ? h[3 * c + 1] : 1.0;
      hz[e] = e < nc ? h[3 * c + 2] : 1.0;
#pragma unroll
      for (int b = 0; b < NPC; ++b) node[e][b] = e < nc ? cells[c * NPC + b] : -1;
    }
    for (int32_t q = q0; q < q1; ++q) {
      const int32_t j = col[q];
      const int32_t jc = perm ? perm[j] : j;
      double acc = 0.0;
#pragma unroll
      for (int e = 0; e < kMaxRowCells; ++e) {
#pragma unroll
        for (int b = 0; b < NPC; ++b) {
          if (node[e][b] == jc) {
            acc = add_rn(acc, mul_rn(kk[e], q1_entry<DIM>(loc[e], b, hx[e], hy[e], hz[e])));
            ++matched;
          }
        }
      }
      if (drow)
        acc = (j == i) ? 1.0 : 0.0;
      else if (isdir && isdir[j])
        acc = 0.0;
      val[q] = acc;
    }
  } else {
    for (int32_t q = q0; q < q1; ++q) {
      const int32_t j = col[q];
      const int32_t jc = perm ? perm[j] : j;
      double acc = 0.0;
      for (int32_t e = c0; e < c1; ++e) {
        const int32_t ce = nlist[e];
        const int64_t c = ce >> 3;
        const int32_t* v = cells + c * NPC;
#pragma unroll
        for (int b = 0; b < NPC; ++b) {
          if (v[b] == jc) {
            const double k = kc ? kc[c] : 1.0;
            acc = add_rn(acc, mul_rn(k, q1_entry<DIM>(ce & 7, b, h[3 * c], h[3 * c + 1], h[3 * c + 2])));
            ++matched;
          }
        }
      }
      if (drow)
        acc = (j == i) ? 1.0 : 0.0;
      else if (isdir && isdir[j])
        acc = 0.0;
      val[q] = acc;
    }
  }
  if (matched != nc * NPC) flags[2] = 1;
}

// One thread per node: r_i = f_i - sum over the cells c of node i (ascending)
// of k_c sum_b K^c_ab u(v_b); Dirichlet rows zero.
template <int DIM>
__global__ __launch_bounds__(kBlock) void k_form_q1(
    int64_t n, const int32_t* __restrict__ nptr, const int32_t* __restrict__ nlist,
    const int32_t* __restrict__ cells, const double* __restrict__ h,
    const double* __restrict__ kc, const uint8_t* __restrict__ isdir,
    const double* __restrict__ u, const double* __restrict__ f, double* __restrict__ r) {
  constexpr int NPC = DIM == 3 ? 8 : 4;
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  if (isdir && isdir[i]) {
    r[i] = 0.0;
    return;
  }
  double acc = f[i];
  for (int32_t e = nptr[i]; e < nptr[i + 1]; ++e) {
    const int32_t ce = nlist[e];
    const int64_t c = ce >> 3;
    const int a = ce & 7;
    const int32_t* v = cells + c * NPC;
    const double hx = h[3 * c], hy = h[3 * c + 1], hz = h[3 * c + 2];
    double s = 0.0;
#pragma unroll
    for (int b = 0; b < NPC; ++b) s = add_rn(s, mul_rn(q1_entry<DIM>(a, b, hx, hy, hz), u[v[b]]));
    acc = add_rn(acc, -mul_rn(kc ? kc[c] : 1.0, s));
  }
  r[i] = acc;
}

// ------------------------------------------------ triquadratic patch elements
// The Q2 element of CGDiscQ23d (src/DofHandler/cgdisc.h:1292) on a patch: the
// 27 patch nodes carry the triquadratic Lagrange basis (BaseQ23d), and for the
// Laplacian on an axis-aligned box the element matrix is a sum of tensor
// products of the 1D quadratic stiffness S = (7,-8,1;-8,16,-8;1,-8,7)/(3L) and
// mass M = (4,2,-1;2,16,2;-1,2,4) L/30 on [0, L], L = 2h (the reference's
// 3x3x3 Gauss rule is exact for these products):
//   E_ab = S_x M_y M_z + M_x S_y M_z + M_x M_y S_z,
// each factor, product and sum rounded separately in that order, as the host
// producer forms it.
__device__ __forceinline__ double q2s(int p, int q, double L) {
  const double s = (p == q) ? (p == 1 ? 16.0 : 7.0) : ((p == 1 || q == 1) ? -8.0 : 1.0);
  return s / mul_rn(3.0, L);
}
__device__ __forceinline__ double q2m(int p, int q, double L) {
  const double m = (p == q) ? (p == 1 ? 16.0 : 4.0) : ((p == 1 || q == 1) ? 2.0 : -1.0);
  return mul_rn(m, L) / 30.0;
}
__device__ __forceinline__ double q2_entry(int a, int b, const double* __restrict__ hs) {
  const double lx = mul_rn(2.0, hs[0]), ly = mul_rn(2.0, hs[1]), lz = mul_rn(2.0, hs[2]);
  const int ax = a % 3, ay = (a / 3) % 3, az = a / 9, bx = b % 3, by = (b / 3) % 3, bz = b / 9;
  const double sx = q2s(ax, bx, lx), sy = q2s(ay, by, ly), sz = q2s(az, bz, lz);
  const double mx = q2m(ax, bx, lx), my = q2m(ay, by, ly), mz = q2m(az, bz, lz);
  const double t1 = mul_rn(mul_rn(sx, my), mz);
  const double t2 = mul_rn(mul_rn(mx, sy), mz);
  const double t3 = mul_rn(mul_rn(mx, my), sz);
  return add_rn(add_rn(t1, t2), t3);
}

// One wavefront per row i (the matrix's numbering; its node is perm[i], or
// i): the row's column keys and its running values sit in LDS; for each patch
// of the node, in ascending order, lane b takes the patch's node b, finds its
// entry by bisection among the keys and adds k_p E_ab there, as the host
// producer adds them (one patch after the other, so each entry receives its
// patches in ascending order); then the Dirichlet rows become unit rows and
// the Dirichlet columns zero.  The keys: the stored columns, with the node's
// position found through iperm (mode 0), or the columns mapped through perm
// (mode 1, the renumbered operators that keep the caller's entry order);
// either must ascend along the row.  A patch node without an entry in the row
// sets flags[2].  Rows hold at most kQ2RowMax entries (checked by the host).
constexpr int kQ2RowMax = 256;
constexpr int kQ2Waves = kBlock / 64;

__global__ __launch_bounds__(kBlock) void k_assemble_q2(
    int64_t n, const int32_t* __restrict__ rowptr, const int32_t* __restrict__ col,
    const int32_t* __restrict__ perm, const int32_t* __restrict__ iperm, int mode,
    const int32_t* __restrict__ pptr,
    const int32_t* __restrict__ plist, const int32_t* __restrict__ nodes,
    const double* __restrict__ hs, const double* __restrict__ kp,
    const uint8_t* __restrict__ isdir, double* __restrict__ val, int* flags) {
  __shared__ int32_t s_key[kQ2Waves][kQ2RowMax];
  __shared__ double s_acc[kQ2Waves][kQ2RowMax];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t i = (int64_t)blockIdx.x * kQ2Waves + w;
  if (i >= n) return;  // the whole wavefront: no block barriers below
  int32_t* key = s_key[w];
  double* acc = s_acc[w];
  const int32_t ic = perm ? perm[i] : (int32_t)i;
  const int32_t q0 = rowptr[i], len = rowptr[i + 1] - q0;
  for (int t = lane; t < len; t += 64) {
    key[t] = (perm && mode != 0) ? perm[col[q0 + t]] : col[q0 + t];
    acc[t] = 0.0;
  }
  // the LDS operations of one wavefront complete in issue order, so the lanes
  // see each other's keys and sums from here on
  __builtin_amdgcn_wave_barrier();
  bool bad = false;
  for (int32_t e = pptr[ic]; e < pptr[ic + 1]; ++e) {
    const int64_t p = plist[e] >> 5;
    const int a = plist[e] & 31;
    if (lane < 27) {
      const int32_t v = nodes[p * 27 + lane];
      const int32_t j = (perm && mode == 0) ? iperm[v] : v;
      int32_t lo = 0, hi = len;  // first key >= j
      while (lo < hi) {
        const int32_t mid = lo + ((hi - lo) >> 1);
        if (key[mid] < j) lo = mid + 1;
        else hi = mid;
      }
      if (lo < len && key[lo] == j)
        acc[lo] = add_rn(acc[lo], mul_rn(kp ? kp[p] : 1.0, q2_entry(a, lane, hs + 3 * p)));
      else
        bad = true;
    }
    __builtin_amdgcn_wave_barrier();
  }
  const bool drow = isdir && isdir[i];
  for (int t = lane; t < len; t += 64) {
    const int32_t c = col[q0 + t];
    double x = acc[t];
    if (drow) x = (c == i) ? 1.0 : 0.0;
    else if (isdir && isdir[c]) x = 0.0;
    val[q0 + t] = x;
  }
  if (bad) flags[2] = 1;
}

// One thread per node: r_i = f_i - sum over the patches p of node i
// (ascending) of k_p sum_b E_ab u(v_b); Dirichlet rows zero.
__global__ __launch_bounds__(kBlock) void k_form_q2(
    int64_t n, const int32_t* __restrict__ pptr, const int32_t* __restrict__ plist,
    const int32_t* __restrict__ nodes, const double* __restrict__ hs,
    const double* __restrict__ kp, const uint8_t* __restrict__ isdir,
    const double* __restrict__ u, const double* __restrict__ f, double* __restrict__ r) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  if (isdir && isdir[i]) {
    r[i] = 0.0;
    return;
  }
  double acc = f[i];
  for (int32_t e = pptr[i]; e < pptr[i + 1]; ++e) {
    const int64_t p = plist[e] >> 5;
    const int a = plist[e] & 31;
    const int32_t* v = nodes + p * 27;
    double s = 0.0;
    for (int b = 0; b < 27; ++b) s = add_rn(s, mul_rn(q2_entry(a, b, hs + 3 * p), u[v[b]]));
    acc = add_rn(acc, -mul_rn(kp ? kp[p] : 1.0, s));
  }
  r[i] = acc;
}

// Q2 operator condensed over the hanging nodes: A_c = T^T A T with T setting
// each hanging value from its parents (CGDisc::LocalToGlobal with
// CondenseHanging, src/DofHandler/cgdisc.h:353-366, in global form, as the
// host producer forms it).  Row i of A_c: a hanging row is an identity row;
// any other row receives, in this order, the entries of row i of A and then
// those of the rows h of the hanging nodes with parent i (weighted w_hi,
// ascending h), each entry (r, j) going to column j, or, for a hanging j, to
// its parents p with the weight w_jp: s * (a_rj * w_jp) and s * a_rj with
// s = 1 or w_hi.  The contributions to one entry are summed in that order
// (the host merges a stable sort by column, which is the same order).  Then
// the Dirichlet rows and columns.  A missing column sets flags[2].
//
// One wavefront per row, the row's columns and running sums in LDS (rows of
// at most kCondMax entries; longer ones, rare, keep them in global memory).  A row that neither holds a hanging column nor has hanging nodes
// attached takes row i of A alone, whose columns are distinct: the lanes add
// its entries at once.  For the other rows (near the refinement interface)
// the lanes find the contributions' targets, and each lane adds, in order,
// the contributions to the sums it owns.
constexpr int kCondMax = 256;
// entries per chunk of a contributing row: 16 keeps the LDS of a wavefront
// at 4.7 KB, so 8 wavefronts per SIMD fit (the kernel waits on memory)
constexpr int kCondChunk = 16;

__global__ __launch_bounds__(kBlock) void k_condense_q2(
    int64_t n, const int32_t* __restrict__ rp, const int32_t* __restrict__ col,
    const double* __restrict__ val, const int32_t* __restrict__ hstart,
    const int32_t* __restrict__ hpar, const double* __restrict__ hw,
    const int32_t* __restrict__ fstart, const int32_t* __restrict__ flist,
    const double* __restrict__ fw, const int32_t* __restrict__ rpc,
    const int32_t* __restrict__ colc, const uint8_t* __restrict__ isdir,
    double* __restrict__ valc, int* flags) {
  __shared__ int32_t s_key[kQ2Waves][kCondMax];
  __shared__ double s_acc[kQ2Waves][kCondMax];
  // a chunk's contributions (target, value): up to 9 per entry, as the host
  // admits one constraint, of at most 9 parents, per hanging node
  __shared__ int32_t s_pos[kQ2Waves][kCondChunk * 9];
  __shared__ double s_x[kQ2Waves][kCondChunk * 9];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t i = (int64_t)blockIdx.x * kQ2Waves + w;
  if (i >= n) return;  // the whole wavefront: no block barriers below
  const int32_t q0 = rpc[i], len = rpc[i + 1] - q0;
  if (hstart[i + 1] > hstart[i]) {
    for (int t = lane; t < len; t += 64) valc[q0 + t] = (colc[q0 + t] == i) ? 1.0 : 0.0;
    return;
  }
  const bool lds = len <= kCondMax;
  const int32_t* keys = lds ? s_key[w] : colc + q0;
  double* acc = lds ? s_acc[w] : valc + q0;
  for (int t = lane; t < len; t += 64) {
    if (lds) s_key[w][t] = colc[q0 + t];
    acc[t] = 0.0;
  }
  // LDS operations of one wavefront complete in issue order; the sums of a
  // long row live in global memory, where each lane's stores must complete
  // before other lanes read them
  if (!lds) __threadfence_block();
  __builtin_amdgcn_wave_barrier();
  bool bad = false;
  auto find = [&](int32_t c) {
    int32_t lo = 0, hi = len;
    while (lo < hi) {
      const int32_t mid = lo + ((hi - lo) >> 1);
      if (keys[mid] < c) lo = mid + 1;
      else hi = mid;
    }
    return (lo < len && keys[lo] == c) ? lo : -1;
  };
  // does row i of A hold a hanging column?
  bool hang_col = false;
  for (int32_t q = rp[i] + lane; q < rp[i + 1]; q += 64) {
    const int32_t j = col[q];
    hang_col = hang_col || hstart[j + 1] > hstart[j];
  }
  const bool simple = !__any(hang_col) && fstart[i + 1] == fstart[i];
  if (simple) {
    for (int32_t q = rp[i] + lane; q < rp[i + 1]; q += 64) {
      const int32_t at = find(col[q]);
      if (at >= 0) acc[at] = add_rn(acc[at], mul_rn(1.0, val[q]));
      else bad = true;
    }
  } else {
    // in chunks of kCondChunk entries of each contributing row: a lane finds
    // the targets of its entry's contributions (one, or the 3 or 9 parents of a
    // hanging column); the chunk's contributions go to a list in sequence
    // order, and each lane adds those whose target it owns (target mod 64),
    // so the contributions to every sum are added in sequence order
    int32_t* f_pos = s_pos[w];
    double* f_x = s_x[w];
    auto segment = [&](int64_t r, double sc) {
      for (int32_t base = rp[r]; base < rp[r + 1]; base += kCondChunk) {
        const int32_t q = base + lane;
        int32_t j = -1, m = 0;
        double v = 0.0;
        if (lane < kCondChunk && q < rp[r + 1]) {
          j = col[q];
          v = val[q];
          m = hstart[j + 1] > hstart[j] ? hstart[j + 1] - hstart[j] : 1;
        }
        // exclusive prefix of m over the lanes
        int32_t inc = m;
        for (int d = 1; d < 64; d <<= 1) {
          const int32_t t = __shfl_up(inc, d, 64);
          if (lane >= d) inc += t;
        }
        const int32_t total = __shfl(inc, 63, 64);
        int32_t off = inc - m;
        if (m > 0) {
          if (hstart[j + 1] > hstart[j]) {
            for (int32_t k = hstart[j]; k < hstart[j + 1]; ++k, ++off) {
              f_pos[off] = find(hpar[k]);
              f_x[off] = mul_rn(sc, mul_rn(v, hw[k]));
            }
          } else {
            f_pos[off] = find(j);
            f_x[off] = mul_rn(sc, v);
          }
        }
        __builtin_amdgcn_wave_barrier();
        for (int32_t f = 0; f < total; ++f) {
          const int32_t at = f_pos[f];
          if (at < 0) bad = true;
          else if ((at & 63) == lane) acc[at] = add_rn(acc[at], f_x[f]);
        }
        __builtin_amdgcn_wave_barrier();
      }
    };
    segment(i, 1.0);
    for (int32_t k = fstart[i]; k < fstart[i + 1]; ++k) segment(flist[k], fw[k]);
  }
  if (!lds) __threadfence_block();
  __builtin_amdgcn_wave_barrier();
  const bool drow = isdir && isdir[i];
  for (int t = lane; t < len; t += 64) {
    const int32_t c = colc[q0 + t];
    double x = acc[t];
    if (drow) x = (c == i) ? 1.0 : 0.0;
    else if (isdir && isdir[c]) x = 0.0;
    valc[q0 + t] = x;
  }
  if (bad) flags[2] = 1;
}

}  // namespace

namespace gh {

int assemble_q1_launch(gh_ctx* ctx, const gh_cells* C, const double* kcell, const uint8_t* isdir,
                       const int32_t* perm, gh_csr* A, double* val64) {
  if (!A->d_flags) {
    if (hipMalloc((void**)&A->d_flags, sizeof(int) * 4) != hipSuccess)
      return fail(GH_ENOMEM, "device assembly: flags");
    GH_HIP(hipMemsetAsync(A->d_flags, 0, sizeof(int) * 4, ctx->stream));
  }
  if (A->n == 0) return GH_OK;
  const dim3 g(grid_for(A->n));
  if (C->dim == 3)
    hipLaunchKernelGGL(k_assemble_q1<3>, g, dim3(kBlock), 0, ctx->stream, A->n, A->rowptr, A->col,
                       perm, C->nptr, C->nlist, C->cells, C->h, kcell, isdir, val64, A->d_flags);
  else
    hipLaunchKernelGGL(k_assemble_q1<2>, g, dim3(kBlock), 0, ctx->stream, A->n, A->rowptr, A->col,
                       perm, C->nptr, C->nlist, C->cells, C->h, kcell, isdir, val64, A->d_flags);
  GH_HIP(hipGetLastError());
  return GH_OK;
}

// How k_assemble_q2 finds a row's entries in A (perm_h: A's row i is node
// perm_h[i]): 0 bisection on the stored columns, 1 on the columns mapped
// through perm, -2 neither ascends (refused), -3 a row longer than
// kQ2RowMax.  Cached in A (one numbering per matrix).
int assemble_q2_mode(gh_csr* A, const std::vector<int32_t>* perm_h) {
  if (A->asm_mode != -1) return A->asm_mode;
  for (int64_t i = 0; i < A->n; ++i)
    if (A->h_rowptr[i + 1] - A->h_rowptr[i] > kQ2RowMax) return A->asm_mode = -3;
  auto ascend = [&](bool mapped) {
    for (int64_t i = 0; i < A->n; ++i)
      for (int32_t q = A->h_rowptr[i] + 1; q < A->h_rowptr[i + 1]; ++q) {
        const int32_t c0 = A->h_col[q - 1], c1 = A->h_col[q];
        if (mapped ? (*perm_h)[c1] <= (*perm_h)[c0] : c1 <= c0) return false;
      }
    return true;
  };
  if (ascend(false)) A->asm_mode = 0;
  else A->asm_mode = (perm_h && ascend(true)) ? 1 : -2;
  return A->asm_mode;
}

int assemble_q2_launch(gh_ctx* ctx, const gh_patches* P, const double* kp, const uint8_t* isdir,
                       const int32_t* perm, const int32_t* iperm, int mode, gh_csr* A,
                       double* val64) {
  if (!A->d_flags) {
    if (hipMalloc((void**)&A->d_flags, sizeof(int) * 4) != hipSuccess)
      return fail(GH_ENOMEM, "device assembly: flags");
    GH_HIP(hipMemsetAsync(A->d_flags, 0, sizeof(int) * 4, ctx->stream));
  }
  if (mode < 0)
    return fail(GH_EINVAL, mode == -3 ? "Q2 assembly: a row has more than 256 entries"
                                      : "Q2 assembly: the matrix's columns do not ascend in every row");
  if (A->n == 0) return GH_OK;
  const int64_t blocks = (A->n + kQ2Waves - 1) / kQ2Waves;
  hipLaunchKernelGGL(k_assemble_q2, dim3((unsigned)blocks), dim3(kBlock), 0, ctx->stream, A->n,
                     A->rowptr, A->col, perm, iperm, mode, P->pptr, P->plist, P->nodes, P->hs,
                     kp, isdir, val64, A->d_flags);
  GH_HIP(hipGetLastError());
  return GH_OK;
}

}  // namespace gh

extern "C" {

int gh_assemble_q1(gh_ctx* ctx, const gh_cells* C, const gh_vec* k, const gh_mask* dirichlet,
                   gh_csr* A) {
  GH_ARG(ctx && C && A, "gh_assemble_q1: NULL argument");
  GH_ARG(!A->host_only && A->ncomp == 1,
         "gh_assemble_q1: needs a scalar matrix with a device layout");
  GH_ARG(A->n == C->n, "gh_assemble_q1: the matrix and the cells have different node counts");
  GH_ARG(!k || (k->prec == GH_F64 && k->size == C->ncells),
         "gh_assemble_q1: k must be an FP64 vector with one entry per cell");
  GH_ARG(!dirichlet || (dirichlet->n == A->n && dirichlet->ncomp == 1),
         "gh_assemble_q1: the Dirichlet mask does not match the matrix");
  const uint8_t* isdir = nullptr;
  if (dirichlet) GH_TRY(mask_node_flags(ctx, const_cast<gh_mask*>(dirichlet), &isdir));
  double* v64 = nullptr;
  GH_TRY(csr_asm_values(ctx, A, &v64));
  GH_TRY(assemble_q1_launch(ctx, C, k ? (const double*)k->d : nullptr, isdir, nullptr, A, v64));
  return csr_asm_commit(ctx, A, "gh_assemble_q1");
}

int gh_form_q1(gh_ctx* ctx, const gh_cells* C, const gh_vec* k, const gh_mask* dirichlet,
               const gh_vec* u, const gh_vec* f, gh_vec* r) {
  GH_ARG(ctx && C && u && f && r, "gh_form_q1: NULL argument");
  GH_ARG(u->prec == GH_F64 && f->prec == GH_F64 && r->prec == GH_F64,
         "gh_form_q1: vectors must be FP64");
  GH_ARG(u->ncomp == 1 && f->ncomp == 1 && r->ncomp == 1 && u->n == C->n && f->n == C->n &&
             r->n == C->n,
         "gh_form_q1: u, f and r must be scalar nodal vectors of the mesh");
  GH_ARG(r->d != u->d, "gh_form_q1: r and u must not alias");
  GH_ARG(!k || (k->prec == GH_F64 && k->size == C->ncells),
         "gh_form_q1: k must be an FP64 vector with one entry per cell");
  GH_ARG(!dirichlet || (dirichlet->n == C->n && dirichlet->ncomp == 1),
         "gh_form_q1: the Dirichlet mask does not match the mesh");
  const uint8_t* isdir = nullptr;
  if (dirichlet) GH_TRY(mask_node_flags(ctx, const_cast<gh_mask*>(dirichlet), &isdir));
  if (C->n == 0) return GH_OK;
  const dim3 g(grid_for(C->n));
  const double* kc = k ? (const double*)k->d : nullptr;
  if (C->dim == 3)
    hipLaunchKernelGGL(k_form_q1<3>, g, dim3(kBlock), 0, ctx->stream, C->n, C->nptr, C->nlist,
                       C->cells, C->h, kc, isdir, (const double*)u->d, (const double*)f->d,
                       (double*)r->d);
  else
    hipLaunchKernelGGL(k_form_q1<2>, g, dim3(kBlock), 0, ctx->stream, C->n, C->nptr, C->nlist,
                       C->cells, C->h, kc, isdir, (const double*)u->d, (const double*)f->d,
                       (double*)r->d);
  GH_HIP(hipGetLastError());
  return GH_OK;
}

int gh_assemble_q2(gh_ctx* ctx, const gh_patches* P, const gh_vec* k, const gh_mask* dirichlet,
                   gh_csr* A) {
  GH_ARG(ctx && P && A, "gh_assemble_q2: NULL argument");
  GH_ARG(P->dim == 3, "gh_assemble_q2: triquadratic patch elements are three-dimensional");
  GH_ARG(!A->host_only && A->ncomp == 1,
         "gh_assemble_q2: needs a scalar matrix with a device layout");
  GH_ARG(A->n == P->nnodes, "gh_assemble_q2: the matrix and the patches have different node counts");
  GH_ARG(!k || (k->prec == GH_F64 && k->size == P->npatches),
         "gh_assemble_q2: k must be an FP64 vector with one entry per patch");
  GH_ARG(!dirichlet || (dirichlet->n == A->n && dirichlet->ncomp == 1),
         "gh_assemble_q2: the Dirichlet mask does not match the matrix");
  const uint8_t* isdir = nullptr;
  if (dirichlet) GH_TRY(mask_node_flags(ctx, const_cast<gh_mask*>(dirichlet), &isdir));
  double* v64 = nullptr;
  GH_TRY(csr_asm_values(ctx, A, &v64));
  GH_TRY(assemble_q2_launch(ctx, P, k ? (const double*)k->d : nullptr, isdir, nullptr, nullptr,
                            assemble_q2_mode(A, nullptr), A, v64));
  return csr_asm_commit(ctx, A, "gh_assemble_q2");
}

int gh_condense_q2(gh_ctx* ctx, int64_t nedges, const int32_t* edges, int64_t nfaces,
                   const int32_t* faces, const gh_csr* A, const gh_mask* dirichlet, gh_csr* Ac) {
  GH_ARG(ctx && A && Ac && nedges >= 0 && nfaces >= 0 && (nedges == 0 || edges) &&
             (nfaces == 0 || faces),
         "gh_condense_q2: bad arguments");
  GH_ARG(!A->host_only && A->ncomp == 1 && A->prec == GH_F64,
         "gh_condense_q2: A must be a scalar FP64 matrix with a device layout");
  GH_ARG(!Ac->host_only && Ac->ncomp == 1 && Ac->n == A->n,
         "gh_condense_q2: Ac must be a scalar matrix of A's size with a device layout");
  GH_ARG(!dirichlet || (dirichlet->n == A->n && dirichlet->ncomp == 1),
         "gh_condense_q2: the Dirichlet mask does not match the matrix");
  if (Ac->cols_ascend < 0) {
    Ac->cols_ascend = 1;
    for (int64_t i = 0; i < Ac->n && Ac->cols_ascend; ++i)
      for (int32_t q = Ac->h_rowptr[i] + 1; q < Ac->h_rowptr[i + 1]; ++q)
        if (Ac->h_col[q] <= Ac->h_col[q - 1]) {
          Ac->cols_ascend = 0;
          break;
        }
  }
  GH_ARG(Ac->cols_ascend == 1, "gh_condense_q2: Ac's columns must ascend in every row");
  const int64_t n = A->n;
  // hanging node -> (parent, weight): the edge constraints' parents with
  // (3/8, 3/4, -1/8), then the face constraints' with the products
  const double w1[3] = {0.375, 0.75, -0.125};
  std::vector<int32_t> hstart(n + 1, 0);
  for (int64_t k = 0; k < nedges; ++k) {
    const int32_t h = edges[4 * k];
    GH_ARG(h >= 0 && h < n, "gh_condense_q2: edge constraint node out of range");
    GH_ARG(hstart[h + 1] == 0, "gh_condense_q2: a node is hanging in two constraints");
    hstart[h + 1] += 3;
  }
  for (int64_t k = 0; k < nfaces; ++k) {
    const int32_t h = faces[10 * k];
    GH_ARG(h >= 0 && h < n, "gh_condense_q2: face constraint node out of range");
    GH_ARG(hstart[h + 1] == 0, "gh_condense_q2: a node is hanging in two constraints");
    hstart[h + 1] += 9;
  }
  for (int64_t i = 0; i < n; ++i) hstart[i + 1] += hstart[i];
  std::vector<int32_t> hpar(hstart[n]);
  std::vector<double> hw(hstart[n]);
  {
    std::vector<int32_t> fill(hstart.begin(), hstart.end() - 1);
    for (int64_t k = 0; k < nedges; ++k)
      for (int t = 0; t < 3; ++t) {
        const int32_t q = fill[edges[4 * k]]++;
        hpar[q] = edges[4 * k + 1 + t];
        hw[q] = w1[t];
      }
    for (int64_t k = 0; k < nfaces; ++k)
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
          const int32_t q = fill[faces[10 * k]]++;
          hpar[q] = faces[10 * k + 1 + 3 * a + b];
          hw[q] = w1[a] * w1[b];
        }
  }
  for (int32_t p : hpar) GH_ARG(p >= 0 && p < n, "gh_condense_q2: constraint parent out of range");
  // parent -> hanging nodes that feed it, ascending hanging node
  std::vector<int32_t> fstart(n + 1, 0), flist(hpar.size());
  std::vector<double> fw(hpar.size());
  for (int32_t p : hpar) fstart[p + 1]++;
  for (int64_t i = 0; i < n; ++i) fstart[i + 1] += fstart[i];
  {
    std::vector<int32_t> fill(fstart.begin(), fstart.end() - 1);
    for (int64_t h = 0; h < n; ++h)
      for (int32_t k = hstart[h]; k < hstart[h + 1]; ++k) {
        const int32_t q = fill[hpar[k]]++;
        flist[q] = (int32_t)h;
        fw[q] = hw[k];
      }
  }
  const uint8_t* isdir = nullptr;
  if (dirichlet) GH_TRY(mask_node_flags(ctx, const_cast<gh_mask*>(dirichlet), &isdir));
  double* vc = nullptr;
  GH_TRY(csr_asm_values(ctx, Ac, &vc));
  if (!Ac->d_flags) {
    if (hipMalloc((void**)&Ac->d_flags, sizeof(int) * 4) != hipSuccess)
      return fail(GH_ENOMEM, "gh_condense_q2: flags");
    GH_HIP(hipMemsetAsync(Ac->d_flags, 0, sizeof(int) * 4, ctx->stream));
  }
  int32_t *d_hs = nullptr, *d_hp = nullptr, *d_fs = nullptr, *d_fl = nullptr;
  double *d_hw = nullptr, *d_fw = nullptr;
  auto up = [&](void** d, const void* h, size_t bytes) {
    if (hipMalloc(d, std::max<size_t>(bytes, 8)) != hipSuccess) return false;
    return bytes == 0 || hipMemcpy(*d, h, bytes, hipMemcpyHostToDevice) == hipSuccess;
  };
  int rc = GH_OK;
  if (!up((void**)&d_hs, hstart.data(), 4 * hstart.size()) ||
      !up((void**)&d_hp, hpar.data(), 4 * hpar.size()) ||
      !up((void**)&d_hw, hw.data(), 8 * hw.size()) ||
      !up((void**)&d_fs, fstart.data(), 4 * fstart.size()) ||
      !up((void**)&d_fl, flist.data(), 4 * flist.size()) ||
      !up((void**)&d_fw, fw.data(), 8 * fw.size()))
    rc = fail(GH_ENOMEM, "gh_condense_q2: constraint tables");
  if (!rc && n > 0) {
    const int64_t blocks = (n + kQ2Waves - 1) / kQ2Waves;
    hipLaunchKernelGGL(k_condense_q2, dim3((unsigned)blocks), dim3(kBlock), 0, ctx->stream, n,
                       A->rowptr, A->col, (const double*)A->val, d_hs, d_hp, d_hw, d_fs, d_fl,
                       d_fw, Ac->rowptr, Ac->col, isdir, vc, Ac->d_flags);
    if (hipGetLastError() != hipSuccess) rc = fail(GH_EHIP, "gh_condense_q2: launch failed");
  }
  if (!rc) rc = csr_asm_commit(ctx, Ac, "gh_condense_q2");  // synchronises
  else (void)hipStreamSynchronize(ctx->stream);
  for (void* d : {(void*)d_hs, (void*)d_hp, (void*)d_hw, (void*)d_fs, (void*)d_fl, (void*)d_fw})
    if (d) (void)hipFree(d);
  return rc;
}

int gh_form_q2(gh_ctx* ctx, const gh_patches* P, const gh_vec* k, const gh_mask* dirichlet,
               const gh_vec* u, const gh_vec* f, gh_vec* r) {
  GH_ARG(ctx && P && u && f && r, "gh_form_q2: NULL argument");
  GH_ARG(P->dim == 3, "gh_form_q2: triquadratic patch elements are three-dimensional");
  GH_ARG(u->prec == GH_F64 && f->prec == GH_F64 && r->prec == GH_F64,
         "gh_form_q2: vectors must be FP64");
  GH_ARG(u->ncomp == 1 && f->ncomp == 1 && r->ncomp == 1 && u->n == P->nnodes &&
             f->n == P->nnodes && r->n == P->nnodes,
         "gh_form_q2: u, f and r must be scalar nodal vectors of the patches' mesh");
  GH_ARG(r->d != u->d, "gh_form_q2: r and u must not alias");
  GH_ARG(!k || (k->prec == GH_F64 && k->size == P->npatches),
         "gh_form_q2: k must be an FP64 vector with one entry per patch");
  GH_ARG(!dirichlet || (dirichlet->n == P->nnodes && dirichlet->ncomp == 1),
         "gh_form_q2: the Dirichlet mask does not match the mesh");
  const uint8_t* isdir = nullptr;
  if (dirichlet) GH_TRY(mask_node_flags(ctx, const_cast<gh_mask*>(dirichlet), &isdir));
  if (P->nnodes == 0) return GH_OK;
  hipLaunchKernelGGL(k_form_q2, dim3(grid_for(P->nnodes)), dim3(kBlock), 0, ctx->stream,
                     P->nnodes, P->pptr, P->plist, P->nodes, P->hs,
                     k ? (const double*)k->d : nullptr, isdir, (const double*)u->d,
                     (const double*)f->d, (double*)r->d);
  GH_HIP(hipGetLastError());
  return GH_OK;
}

int gh_csr_download_values(gh_ctx* ctx, const gh_csr* A, double* val) {
  GH_ARG(ctx && A && (val || A->nnz == 0), "gh_csr_download_values: NULL argument");
  GH_ARG(!A->host_only, "gh_csr_download_values: the matrix has no device layout");
  const int64_t m = A->nnz * (int64_t)A->ncomp * A->ncomp;
  if (m == 0) return GH_OK;
  if (A->prec == GH_F64) {
    GH_HIP(hipMemcpyAsync(val, A->val, sizeof(double) * m, hipMemcpyDeviceToHost, ctx->stream));
    GH_HIP(hipStreamSynchronize(ctx->stream));
    return GH_OK;
  }
  double* tmp = nullptr;
  if (hipMalloc((void**)&tmp, sizeof(double) * m) != hipSuccess)
    return fail(GH_ENOMEM, "gh_csr_download_values: staging buffer");
  const unsigned g = (unsigned)std::min<int64_t>((m + kBlock - 1) / kBlock, 8192);
  if (A->prec == GH_F32)
    hipLaunchKernelGGL((k_convert<double, float>), dim3(g), dim3(kBlock), 0, ctx->stream, m, tmp,
                       (const float*)A->val);
  else
    hipLaunchKernelGGL((k_convert<double, __half>), dim3(g), dim3(kBlock), 0, ctx->stream, m, tmp,
                       (const __half*)A->val);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(val, tmp, sizeof(double) * m, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  (void)hipFree(tmp);
  if (e != hipSuccess) return fail(GH_EHIP, "gh_csr_download_values: %s", hipGetErrorString(e));
  return GH_OK;
}

}  // extern "C"
