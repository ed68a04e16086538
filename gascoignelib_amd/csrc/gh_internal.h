// Internal definitions of the gascoigne_hip library (gfx950 only).
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>

#include <algorithm>
#include <atomic>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>
#include <thread>
#include <vector>

#include "gascoigne_hip.h"

// Vanka patch smoother: patch LU factors and the node -> patch lists
struct gh_vanka {
  gh_ctx* ctx = nullptr;
  int B = 1, psize = 0, m = 0;
  int prec = GH_F64;          // storage of the factors and of the local solves
  int64_t n = 0, npatches = 0;
  int32_t* nodes = nullptr;   // npatches * psize
  void* lu = nullptr;         // npatches * m * m, column-major, L unit lower
  int32_t* piv = nullptr;     // npatches * m: row exchanges, then (after set-up) the
                              // global entry gathered into each factor row
  int32_t* nptr = nullptr;    // n + 1
  int32_t* nent = nullptr;    // patch * psize + local node, ascending patch per node
  void* xp = nullptr;         // npatches * m local solutions
};

// Q1 box cells for coefficient gradients
struct gh_cells {
  gh_ctx* ctx = nullptr;
  int dim = 3;
  int64_t n = 0;        // nodes
  int64_t ncells = 0;
  int32_t* cells = nullptr;  // ncells * (4 | 8) node indices, local corner order
  double* h = nullptr;       // ncells * 3 box edge lengths (hz = 1 in 2D)
  // node -> cells: the cells of node i are nlist[nptr[i] .. nptr[i+1]), each
  // entry cell * 8 + local corner, in ascending cell order
  int32_t* nptr = nullptr;
  int32_t* nlist = nullptr;
};

// Patches of a refined mesh: the 2^dim children of one coarser cell, 3^dim
// nodes each in lexicographic order (x fastest), axis-aligned boxes with
// midpoint middle nodes.  Used by the DWR indicators (the Q2 interpolation of
// the adjoint) and by the device assembly of triquadratic patch elements.
struct gh_patches {
  gh_ctx* ctx = nullptr;
  int dim = 3;
  int64_t nnodes = 0, npatches = 0;
  int32_t* nodes = nullptr;  // npatches * 3^dim
  double* h = nullptr;       // npatches * dim: half the patch extent per axis
  double* hs = nullptr;      // npatches * 3: node 0 to nodes 1, 3, 9 along x, y, z
  // node -> patches: the patches of node i are plist[pptr[i] .. pptr[i+1]),
  // each entry patch * 32 + local node, in ascending patch order
  int32_t* pptr = nullptr;
  int32_t* plist = nullptr;
};

namespace gh {

// ---------------------------------------------------------------- errors
void set_error(const char* fmt, ...);
int fail(int code, const char* fmt, ...);

#define GH_HIP(call)                                                              \
  do {                                                                            \
    hipError_t e_ = (call);                                                       \
    if (e_ != hipSuccess)                                                         \
      return ::gh::fail(GH_EHIP, "%s:%d %s: %s", __FILE__, __LINE__, #call,       \
                        hipGetErrorString(e_));                                   \
  } while (0)

#define GH_TRY(expr)        \
  do {                      \
    int rc_ = (expr);       \
    if (rc_ != GH_OK) return rc_; \
  } while (0)

#define GH_ARG(cond, msg)                                  \
  do {                                                     \
    if (!(cond)) return ::gh::fail(GH_EINVAL, "%s", msg); \
  } while (0)

// launch geometry
constexpr int kBlock = 256;       // 4 wavefronts of 64
constexpr int kRedBlocks = 1024;  // fixed grid of every reduction: deterministic sums

inline unsigned grid_for(int64_t work, int per_block = kBlock) {
  int64_t g = (work + per_block - 1) / per_block;
  if (g < 1) g = 1;
  return (unsigned)g;
}

// number of device scalar slots for reduction results
constexpr int kScalars = 512;

constexpr int kSlice = 64;

struct Sell {
  int64_t n = 0;          // rows
  int64_t nslices = 0;
  int64_t slots = 0;      // padded entries
  int64_t stored = 0;     // entries kept (valued operators drop exact zeros)
  int64_t narrow = 0;     // slices with 16-bit column offsets
  int maxlen = 0;         // longest row (entries kept)
  uint32_t* sptr = nullptr;  // nslices+1 slot offsets
  uint16_t* rlen = nullptr;  // entries per row
  // valued operators, per slice one of three column encodings:
  //   pattern (cptr bit 30): every row follows a shared relative pattern,
  //     rlen[row] holds its id, columns are row + ptab[pstart[id] + k]; with
  //     bit 29 all 64 rows share one pattern, whose id is cbase[s];
  //   16-bit offsets from cbase[s] (cptr = start in 16-bit units);
  //   32-bit indices (cptr bit 31, start in 16-bit units, 4-byte aligned)
  uint16_t* cidx = nullptr;
  uint32_t* cptr = nullptr;
  uint32_t* cbase = nullptr;
  int32_t* ptab = nullptr;
  int32_t* pstart = nullptr;
  int64_t npat = 0, ntab = 0;
  int64_t patterned = 0;     // slices in pattern mode
  int64_t col_bytes = 0;     // column bytes a sweep reads (stored entries only)
  // row run by each slot (valued operators whose rows were grouped by
  // pattern; nullptr: slot = row); rlen and the slot arrays are in slot order
  int32_t* perm = nullptr;
  // packed-weight operators: column | weight code << 29, one per slot
  uint32_t* col = nullptr;
  void* val = nullptr;       // values in `prec`, nullptr for packed-weight operators
  int prec = GH_F64;
};

// Block matrices (B = 2, 3, 4) as component rows: row t = i*B + r of block
// row i lists, in the block row's column order, the blocks whose row r is not
// all zero, each entry = (block column, the B values of row r).  Component
// rows are sorted by length (descending, stable) inside windows of
// kBSellWindow rows and cut into slices of 64; one lane runs one component
// row.  Entry k of lane l of slice s: column at (sptr[s] + k) * 64 + l, value
// c at ((sptr[s] + k) * B + c) * 64 + l (each value load of a wavefront reads
// 64 consecutive elements).
constexpr int64_t kBSellWindow = 4096;
struct BSell {
  int B = 0;              // 0: not built (the CSR kernels run)
  int64_t nrows = 0;      // n * B component rows
  int64_t nslices = 0;
  int64_t slots = 0;      // padded entries (multiple of 64)
  int64_t stored = 0;     // entries kept
  uint32_t* sptr = nullptr;  // nslices+1, in units of 64 entries
  int32_t* row = nullptr;    // nslices*64 component row of each lane (-1: none)
  uint16_t* rlen = nullptr;  // nslices*64 entries of each lane
  int32_t* col = nullptr;    // slots block columns
  void* val = nullptr;       // slots*B values in `prec`
  int prec = GH_F64;
};
int bsell_build(gh_ctx* ctx, int64_t n, int B, const int32_t* rowptr, const int32_t* col,
                const double* vals, int prec, BSell* S);
void bsell_free(BSell* S);

// Host-side layout from CSR.  With vals (double) the operator is valued: exact
// zeros are dropped (each row keeps its other entries in order, so every row
// sum takes the same values in the same order) and values are stored as
// `prec`.  Without vals, col holds packed-weight entries.
int sell_build(gh_ctx* ctx, int64_t n, const int32_t* rowptr, const uint32_t* col,
               const double* vals, int prec, Sell* S);
void sell_free(Sell* S);

}  // namespace gh

struct gh_ctx {
  // the context lives until gh_ctx_destroy was called and the last object
  // created on it is destroyed
  std::atomic<int> refs{1};
  int device = 0;
  hipStream_t stream = nullptr;
  double* partials = nullptr;      // kScalars * kRedBlocks block partials
  double* scal = nullptr;          // device result slots
  double* host_scal = nullptr;     // pinned mirror of scal
  int num_cu = 256;
  // device address of this device's gh_jacobi_noalloc scratch (nullptr when
  // the lookup failed at creation)
  void* noalloc_scratch = nullptr;
  // rocBLAS handle on `stream` for the direct coarse factorisation (created
  // on first use, destroyed with the context)
  void* blas = nullptr;
};

struct gh_vec {
  gh_ctx* ctx = nullptr;
  int64_t n = 0;      // nodes
  int ncomp = 1;
  int prec = GH_F64;  // GH_F64 or GH_F32
  int64_t size = 0;   // n * ncomp
  void* d = nullptr;
  bool owned = true;
};

struct gh_csr {
  gh_ctx* ctx = nullptr;
  int64_t n = 0;       // block rows
  int ncomp = 1;
  int64_t nnz = 0;     // blocks
  int prec = GH_F64;   // value storage
  int32_t* rowptr = nullptr;
  int32_t* col = nullptr;
  void* val = nullptr;   // nnz*B*B in prec (double / float / half)
  void* dinv = nullptr;  // n*B*B inverse diagonal blocks: double (F64) or float
  gh::Sell sell;         // SELL-64 copy used by the scalar (ncomp == 1) kernels
  gh::BSell bsell;       // component-row copy used by the block (ncomp 2..4) products
  // host copies kept for derived structures (transpose, coarse direct solve)
  std::vector<int32_t> h_rowptr, h_col;
  std::vector<double> h_val;
  // created by gh_csr_create_host: the host copies only, no device layout;
  // accepted by gh_mg_create_ordered (which builds its own device matrices)
  // and as the finest-level handle of such a solver, refused by every device
  // operation
  bool host_only = false;
  // transposed structure for gh_spmv_t (built on first use)
  int32_t* t_rowptr = nullptr;
  int32_t* t_col = nullptr;
  int32_t* t_pos = nullptr;  // position of the transposed block in val
  // device assembly (scalar matrices): the host copy h_val keeps the values of
  // creation only until the first device assembly; constructors that read it
  // refuse the matrix afterwards
  bool h_val_ok = true;
  int32_t* d_diag = nullptr;   // position of each row's diagonal entry
  uint32_t* d_kept = nullptr;  // bit q: CSR entry q is stored in the SELL layout
  double* d_val64 = nullptr;   // FP64 CSR values of a reduced-precision matrix
  int* d_flags = nullptr;      // [0] singular diagonal, [1] a left-out entry became nonzero,
                               // [2] an element coupling has no entry in the pattern
  int asm_mode = -1;           // how the Q2 device assembly finds entries (found on first use)
  int cols_ascend = -1;        // 1: every row's columns ascend (found on first use)
};

struct gh_transfer {
  gh_ctx* ctx = nullptr;
  int64_t ncoarse = 0, nfine = 0;
  // restriction: coarse rows; columns are fine nodes, weight code in the top
  // 3 bits (w = 2^-code), sorted ascending (MgInterpolatorMatrix STcoarse)
  int32_t* r_row = nullptr;
  uint32_t* r_col = nullptr;
  // prolongation: fine rows; columns are coarse parents (STfine)
  int32_t* p_row = nullptr;
  uint32_t* p_col = nullptr;
  std::vector<int32_t> h_rrow, h_prow;
  std::vector<uint32_t> h_rcol, h_pcol;
  gh::Sell r_sell, p_sell;  // SELL-64 copies for scalar vectors
};

struct gh_hn {
  gh_ctx* ctx = nullptr;
  int kind = 13;
  int64_t n = 0;
  int64_t ne = 0, nf = 0;
  int32_t* edges = nullptr;  // ne x 4
  int32_t* faces = nullptr;  // nf x 10
  // distribute in gather form: per parent, contributions in reference order
  int64_t np = 0;
  int32_t* d_parent = nullptr;  // np parents
  int32_t* d_row = nullptr;     // np+1
  int32_t* d_src = nullptr;     // hanging node per contribution
  double* d_w = nullptr;        // weight per contribution
  int32_t* hanging = nullptr;   // ne+nf hanging nodes
  std::vector<int32_t> h_edges, h_faces;
};

struct gh_mask {
  gh_ctx* ctx = nullptr;
  int64_t n = 0;
  int ncomp = 1;
  int64_t nidx = 0;
  int ncomps = 0;
  int32_t* idx = nullptr;      // nidx * ncomps flat entries (node*ncomp+comp)
  std::vector<int32_t> h_nodes, h_comps;
  uint8_t* d_nodeflag = nullptr;  // per node 1 if listed (device assembly; built on first use)
};

namespace gh {
// fn(begin, end) over [0, n) in contiguous chunks of at least min_chunk
// items on up to 16 host threads (the calling thread takes the first); the
// chunks write disjoint ranges, so the result does not depend on the thread
// count.
template <class F>
void parallel_for(int64_t n, int64_t min_chunk, F&& fn) {
  const int64_t hw = std::max<int64_t>(1, (int64_t)std::thread::hardware_concurrency());
  const int T = (int)std::min<int64_t>(std::min<int64_t>(16, hw),
                                       std::max<int64_t>(1, n / std::max<int64_t>(1, min_chunk)));
  if (T <= 1) {
    fn((int64_t)0, n);
    return;
  }
  const int64_t step = (n + T - 1) / T;
  std::vector<std::thread> pool;
  for (int t = 1; t < T; ++t) {
    const int64_t a = t * step, b = std::min(n, a + step);
    if (a < b) pool.emplace_back([&fn, a, b] { fn(a, b); });
  }
  fn((int64_t)0, std::min(n, step));
  for (std::thread& th : pool) th.join();
}
// rows of an operator: chunks of at least 32768 rows
template <class F>
void parallel_rows(int64_t n, F&& fn) {
  parallel_for(n, (int64_t)1 << 15, fn);
}

void ctx_retain(gh_ctx* ctx);
void ctx_release(gh_ctx* ctx);
// vector helpers shared by the solver translation unit
int vec_alloc(gh_ctx* ctx, int64_t n, int ncomp, int prec, gh_vec* v);
void vec_free(gh_vec* v);
int reduce_dot(gh_ctx* ctx, int prec, int64_t size, const void* x, const void* y, int slot);
int reduce_norm_l8(gh_ctx* ctx, int prec, int64_t size, const void* x, int slot);
int fetch_scalars(gh_ctx* ctx, int first, int count, double* out);
int csr_spmv(gh_ctx* ctx, const gh_csr* A, int yprec, void* y, int xprec, const void* x,
             double d, double b);
int csr_residual(gh_ctx* ctx, const gh_csr* A, int prec, void* r, const void* x, const void* b);
int csr_spmv_x32(gh_ctx* ctx, const gh_csr* A, double* y, const float* x);
int vanka_create(gh_ctx* ctx, const gh_csr* A, int64_t npatches, int psize, const int32_t* nodes,
                 int prec, gh_vanka** out);
int vanka_solve(gh_ctx* ctx, const gh_vanka* V, int prec, void* x);
// the inverse of an FP64 matrix (scalar or node blocks) as a dense row-major
// m x m array on the device (m = n * ncomp), factorised on the device
int dense_inverse_device(gh_ctx* ctx, const gh_csr* A, double* inv);
void direct_ctx_free(gh_ctx* ctx);
// SparseBlockIlu::solve of `in` (FP64 or FP32 storage): mode 0 out = h,
// 1 out += omega h, 2 out = omega h (`in` and `out` may alias in mode 0)
int ilu_apply(gh_ctx* ctx, const gh_ilu* I, int prec, const void* in, void* out, int mode,
              double omega);
// host transpose of a node-block CSR (blocks transposed too); rows of the
// result list their columns ascending
void csr_transpose_host(int64_t n, int B, const int32_t* rowptr, const int32_t* col,
                        const double* val, std::vector<int32_t>& rpT, std::vector<int32_t>& clT,
                        std::vector<double>& vlT);
int csr_jacobi_sweep(gh_ctx* ctx, const gh_csr* A, int prec, void* xnew, const void* x,
                     const void* b, double omega, int smoother, bool x_zero);
int vec_axpby(gh_ctx* ctx, int prec, int64_t size, void* y, double a, double d, const void* x);
int vec_scale_to(gh_ctx* ctx, int prec, int64_t size, void* y, double s, const void* x);
int vec_convert(gh_ctx* ctx, int64_t size, int dprec, void* dst, int sprec, const void* src);
int vec_fill_zero(gh_ctx* ctx, int prec, int64_t size, void* x);
int transfer_restrict(gh_ctx* ctx, const gh_transfer* T, int ncomp, int prec, void* uc,
                      const void* uf, const int32_t* zero_rows_mask);
int transfer_prolongate_add(gh_ctx* ctx, const gh_transfer* T, int ncomp, int prec, void* uf,
                            const void* uc);
int hn_apply(gh_ctx* ctx, const gh_hn* H, int ncomp, int prec, void* x, int op);
int mask_apply(gh_ctx* ctx, const gh_mask* M, int prec, void* x);
int mgs_step(gh_ctx* ctx, int64_t size, double* vv, const double* prev, int dslot,
             const double* next, int oslot);
int finalize_slots(gh_ctx* ctx, int first, int count, bool max);
// classical Gram-Schmidt: scal[slot0 + k] = <w, basis[k]> (k < dim); then
// w -= sum_k scal[hslot0 + k] basis[k] and scal[oslot] = <w, w>
int cgs_dots(gh_ctx* ctx, int64_t size, const double* w, const double* const* dbasis, int dim,
             int slot0);
int cgs_update(gh_ctx* ctx, int64_t size, double* w, const double* const* dbasis, int dim,
               int hslot0, int oslot);
int csr_create_impl(gh_ctx* ctx, int64_t n, int ncomp, const int32_t* rowptr, const int32_t* col,
                    const double* val, int prec, bool check_sorted, gh_csr** out);
// Device assembly of scalar matrices.  csr_asm_values: the FP64 CSR value
// array an assembly kernel writes (the matrix's own values in FP64, else a
// buffer allocated on first use).  csr_asm_commit: after that array was
// written on the stream, derive everything else from it (inverse diagonal,
// reduced-precision values, the SELL value stream) and report a singular
// diagonal, a left-out entry that became nonzero or a coupling missing from
// the pattern (flag [2], set by the assembly kernel); synchronises.
int csr_asm_values(gh_ctx* ctx, gh_csr* A, double** val64);
int csr_asm_commit(gh_ctx* ctx, gh_csr* A, const char* who);
// per-node flags of a Dirichlet mask (built on first use)
int mask_node_flags(gh_ctx* ctx, gh_mask* M, const uint8_t** out);
// Q1 assembly of -div(k grad u) on box cells into the FP64 CSR values of A
// (rows and columns in A's numbering; perm[i] = the cells' node of row i,
// nullptr: the same numbering); isdir per row of A (nullptr: none)
int assemble_q1_launch(gh_ctx* ctx, const gh_cells* C, const double* kcell, const uint8_t* isdir,
                       const int32_t* perm, gh_csr* A, double* val64);
// the same for the triquadratic patch elements P (one coefficient per patch);
// mode from assemble_q2_mode (perm_h: the host copy of perm, nullptr: none)
int assemble_q2_mode(gh_csr* A, const std::vector<int32_t>* perm_h);
int assemble_q2_launch(gh_ctx* ctx, const gh_patches* P, const double* kpatch, const uint8_t* isdir,
                       const int32_t* perm, const int32_t* iperm, int mode, gh_csr* A,
                       double* val64);
int transfer_from_csr(gh_ctx* ctx, int64_t ncoarse, int64_t nfine, std::vector<int32_t> rrow,
                      std::vector<uint32_t> rcol, std::vector<int32_t> prow,
                      std::vector<uint32_t> pcol, gh_transfer** out);
// gather: dst[i] = src[perm[i]];  scatter: dst[perm[i]] = src[i]  (node-wise)
int vec_permute(gh_ctx* ctx, int64_t n, int ncomp, const int32_t* perm, int dprec, void* dst,
                int sprec, const void* src, bool gather);
}  // namespace gh
