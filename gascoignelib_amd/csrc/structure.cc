// Host-side structure producer (see include/gh_structure.h): .inp reader,
// nested uniform refinement of quad/hex meshes, mesh-transfer maps, Dirichlet
// lists, CSR structure and Q1 assembly for -div(k grad u).
#include "gh_structure.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <unordered_map>
#include <vector>

namespace {

thread_local std::string g_err;

int fail(const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return -1;
}

struct Mesh {
  int dim = 3;
  int npc = 8;                   // nodes per cell
  std::vector<double> X;         // n * 3
  std::vector<int32_t> cells;    // ncells * npc
  std::vector<int32_t> bnd;      // nb * (dim == 3 ? 4 : 2)
  std::vector<int32_t> bcol;     // nb colours
  // transfer from the next coarser mesh (empty on the coarsest)
  int64_t ncoarse = 0;
  std::vector<int32_t> zweier, vierer, achter;
  int64_t n() const { return (int64_t)X.size() / 3; }
  int64_t ncells() const { return (int64_t)cells.size() / npc; }
};

// parametric corner of local vertex m (Gascoigne's hex / quad orientation:
// 0-3 counter-clockwise at z = 0, 4-7 above them)
const int kHexCorner[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                              {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
const int kQuadCorner[4][2] = {{0, 0}, {1, 0}, {1, 1}, {0, 1}};

int read_inp(const char* path, int dim, Mesh& M) {
  std::ifstream in(path);
  if (!in) return fail("cannot open %s", path);
  int64_t nn, ne;
  int d1, d2, d3;
  in >> nn >> ne >> d1 >> d2 >> d3;
  if (!in || nn <= 0 || ne <= 0) return fail("%s: bad header", path);
  M.dim = dim;
  M.npc = dim == 3 ? 8 : 4;
  M.X.assign(nn * 3, 0.0);
  std::vector<int64_t> idmap;
  std::unordered_map<int64_t, int32_t> id2idx;
  for (int64_t i = 0; i < nn; ++i) {
    int64_t id;
    double x, y, z;
    in >> id >> x >> y >> z;
    if (!in) return fail("%s: bad node line %lld", path, (long long)i);
    id2idx[id] = (int32_t)i;
    M.X[3 * i] = x;
    M.X[3 * i + 1] = y;
    M.X[3 * i + 2] = dim == 3 ? z : 0.0;
  }
  std::string line;
  std::getline(in, line);
  for (int64_t e = 0; e < ne; ++e) {
    if (!std::getline(in, line)) return fail("%s: missing element %lld", path, (long long)e);
    std::istringstream ls(line);
    int64_t id;
    int mat;
    std::string type;
    ls >> id >> mat >> type;
    std::vector<int64_t> v;
    int64_t t;
    while (ls >> t) v.push_back(t);
    auto map_ids = [&](std::vector<int32_t>& dst) -> bool {
      for (int64_t x : v) {
        auto it = id2idx.find(x);
        if (it == id2idx.end()) return false;
        dst.push_back(it->second);
      }
      return true;
    };
    if ((dim == 3 && type == "hex" && v.size() == 8) || (dim == 2 && type == "quad" && v.size() == 4)) {
      if (!map_ids(M.cells)) return fail("%s: unknown node in element %lld", path, (long long)id);
    } else if ((dim == 3 && type == "quad" && v.size() == 4) ||
               (dim == 2 && type == "line" && v.size() == 2)) {
      if (!map_ids(M.bnd)) return fail("%s: unknown node in element %lld", path, (long long)id);
      M.bcol.push_back(mat);
    } else {
      return fail("%s: element %lld of type '%s' does not fit a %dD mesh", path, (long long)id,
                  type.c_str(), dim);
    }
  }
  return 0;
}

inline uint64_t ekey(int32_t a, int32_t b) {
  if (a > b) std::swap(a, b);
  return ((uint64_t)(uint32_t)a << 32) | (uint32_t)b;
}
// a quad face is identified by its smallest vertex and the vertex diagonal to it
inline uint64_t fkey(const int32_t q[4]) {
  int m = 0;
  for (int i = 1; i < 4; ++i)
    if (q[i] < q[m]) m = i;
  return ekey(q[m], q[(m + 2) & 3]);
}

// one uniform refinement (nested numbering)
int refine(const Mesh& C, Mesh& F) {
  const int dim = C.dim;
  F.dim = dim;
  F.npc = C.npc;
  const int64_t n0 = C.n(), nc = C.ncells();
  F.X = C.X;
  F.ncoarse = n0;
  std::vector<double>& X = F.X;
  // cell midpoints
  for (int64_t c = 0; c < nc; ++c) {
    const int32_t* v = &C.cells[c * C.npc];
    double s[3] = {0, 0, 0};
    for (int m = 0; m < C.npc; ++m)
      for (int d = 0; d < 3; ++d) s[d] += C.X[3 * v[m] + d];
    for (int d = 0; d < 3; ++d) X.push_back(s[d] / C.npc);
    const int32_t id = (int32_t)(n0 + c);
    if (dim == 3) {
      F.achter.push_back(id);
      for (int m = 0; m < 8; ++m) F.achter.push_back(v[m]);
    } else {
      F.vierer.push_back(id);
      for (int m = 0; m < 4; ++m) F.vierer.push_back(v[m]);
    }
  }
  auto corner = [&](const int32_t* v, int i, int j, int k) -> int32_t {
    for (int m = 0; m < C.npc; ++m) {
      const int* p = dim == 3 ? kHexCorner[m] : nullptr;
      if (dim == 3 ? (p[0] == i && p[1] == j && p[2] == k)
                   : (kQuadCorner[m][0] == i && kQuadCorner[m][1] == j))
        return v[m];
    }
    return -1;
  };
  // face midpoints (3D)
  std::unordered_map<uint64_t, int32_t> fmap, emap;
  fmap.reserve(dim == 3 ? nc * 4 : 1);
  emap.reserve(nc * (dim == 3 ? 4 : 3));
  if (dim == 3) {
    for (int64_t c = 0; c < nc; ++c) {
      const int32_t* v = &C.cells[c * 8];
      for (int ax = 0; ax < 3; ++ax)
        for (int side = 0; side < 2; ++side) {
          // face with coordinate ax == side; corners in cyclic order
          int32_t q[4];
          const int uv[4][2] = {{0, 0}, {1, 0}, {1, 1}, {0, 1}};
          for (int t = 0; t < 4; ++t) {
            int p[3];
            p[ax] = side;
            p[(ax + 1) % 3] = uv[t][0];
            p[(ax + 2) % 3] = uv[t][1];
            q[t] = corner(v, p[0], p[1], p[2]);
          }
          const uint64_t key = fkey(q);
          if (fmap.find(key) != fmap.end()) continue;
          const int32_t id = (int32_t)(X.size() / 3);
          fmap.emplace(key, id);
          double s[3] = {0, 0, 0};
          for (int t = 0; t < 4; ++t)
            for (int d = 0; d < 3; ++d) s[d] += C.X[3 * q[t] + d];
          for (int d = 0; d < 3; ++d) X.push_back(s[d] / 4.0);
          F.vierer.push_back(id);
          for (int t = 0; t < 4; ++t) F.vierer.push_back(q[t]);
        }
    }
  }
  // edge midpoints
  for (int64_t c = 0; c < nc; ++c) {
    const int32_t* v = &C.cells[c * C.npc];
    const int nedge = dim == 3 ? 12 : 4;
    for (int e = 0; e < nedge; ++e) {
      int32_t a, b;
      if (dim == 3) {
        const int ax = e / 4, r = e % 4;
        int p[3], q[3];
        p[ax] = 0;
        q[ax] = 1;
        p[(ax + 1) % 3] = q[(ax + 1) % 3] = r & 1;
        p[(ax + 2) % 3] = q[(ax + 2) % 3] = (r >> 1) & 1;
        a = corner(v, p[0], p[1], p[2]);
        b = corner(v, q[0], q[1], q[2]);
      } else {
        a = v[e];
        b = v[(e + 1) % 4];
      }
      const uint64_t key = ekey(a, b);
      if (emap.find(key) != emap.end()) continue;
      const int32_t id = (int32_t)(X.size() / 3);
      emap.emplace(key, id);
      for (int d = 0; d < 3; ++d) X.push_back(0.5 * (C.X[3 * a + d] + C.X[3 * b + d]));
      F.zweier.push_back(id);
      F.zweier.push_back(a);
      F.zweier.push_back(b);
    }
  }
  if ((int64_t)(X.size() / 3) > (int64_t)0x1FFFFFFF) return fail("mesh too large for int32 indices");
  // children
  F.cells.clear();
  F.cells.reserve(nc * C.npc * (dim == 3 ? 8 : 4));
  for (int64_t c = 0; c < nc; ++c) {
    const int32_t* v = &C.cells[c * C.npc];
    if (dim == 3) {
      int32_t G[3][3][3];
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b)
          for (int k = 0; k < 3; ++k) {
            const int odd = (a & 1) + (b & 1) + (k & 1);
            if (odd == 0) {
              G[a][b][k] = corner(v, a / 2, b / 2, k / 2);
            } else if (odd == 3) {
              G[a][b][k] = (int32_t)(n0 + c);
            } else if (odd == 1) {
              int p[3] = {a, b, k}, q[3] = {a, b, k};
              for (int d = 0; d < 3; ++d)
                if (p[d] == 1) {
                  p[d] = 0;
                  q[d] = 2;
                }
              G[a][b][k] = emap.at(ekey(corner(v, p[0] / 2, p[1] / 2, p[2] / 2),
                                        corner(v, q[0] / 2, q[1] / 2, q[2] / 2)));
            } else {
              int ax = (a & 1) == 0 ? 0 : (b & 1) == 0 ? 1 : 2;
              const int side = (ax == 0 ? a : ax == 1 ? b : k) / 2;
              int32_t q[4];
              const int uv[4][2] = {{0, 0}, {1, 0}, {1, 1}, {0, 1}};
              for (int t = 0; t < 4; ++t) {
                int p[3];
                p[ax] = side;
                p[(ax + 1) % 3] = uv[t][0];
                p[(ax + 2) % 3] = uv[t][1];
                q[t] = corner(v, p[0], p[1], p[2]);
              }
              G[a][b][k] = fmap.at(fkey(q));
            }
          }
      for (int k = 0; k < 2; ++k)
        for (int j = 0; j < 2; ++j)
          for (int i = 0; i < 2; ++i)
            for (int m = 0; m < 8; ++m)
              F.cells.push_back(G[i + kHexCorner[m][0]][j + kHexCorner[m][1]][k + kHexCorner[m][2]]);
    } else {
      int32_t G[3][3];
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
          const int odd = (a & 1) + (b & 1);
          if (odd == 0)
            G[a][b] = corner(v, a / 2, b / 2, 0);
          else if (odd == 2)
            G[a][b] = (int32_t)(n0 + c);
          else {
            int p[2] = {a, b}, q[2] = {a, b};
            for (int d = 0; d < 2; ++d)
              if (p[d] == 1) {
                p[d] = 0;
                q[d] = 2;
              }
            G[a][b] = emap.at(ekey(corner(v, p[0] / 2, p[1] / 2, 0), corner(v, q[0] / 2, q[1] / 2, 0)));
          }
        }
      for (int j = 0; j < 2; ++j)
        for (int i = 0; i < 2; ++i)
          for (int m = 0; m < 4; ++m)
            F.cells.push_back(G[i + kQuadCorner[m][0]][j + kQuadCorner[m][1]]);
    }
  }
  // boundary elements
  if (dim == 3) {
    const int64_t nb = (int64_t)C.bnd.size() / 4;
    for (int64_t f = 0; f < nb; ++f) {
      const int32_t* q = &C.bnd[4 * f];
      auto fit = fmap.find(fkey(q));
      if (fit == fmap.end()) return fail("boundary quad %lld is not a cell face", (long long)f);
      const int32_t m = fit->second;
      int32_t e[4];
      for (int t = 0; t < 4; ++t) e[t] = emap.at(ekey(q[t], q[(t + 1) & 3]));
      // four children, each keeping the parent's orientation
      const int32_t ch[4][4] = {{q[0], e[0], m, e[3]}, {e[0], q[1], e[1], m},
                                {m, e[1], q[2], e[2]}, {e[3], m, e[2], q[3]}};
      for (int t = 0; t < 4; ++t) {
        F.bnd.insert(F.bnd.end(), ch[t], ch[t] + 4);
        F.bcol.push_back(C.bcol[f]);
      }
    }
  } else {
    const int64_t nb = (int64_t)C.bnd.size() / 2;
    for (int64_t f = 0; f < nb; ++f) {
      const int32_t a = C.bnd[2 * f], b = C.bnd[2 * f + 1];
      auto it = emap.find(ekey(a, b));
      if (it == emap.end()) return fail("boundary line %lld is not a cell edge", (long long)f);
      F.bnd.insert(F.bnd.end(), {a, it->second, it->second, b});
      F.bcol.push_back(C.bcol[f]);
      F.bcol.push_back(C.bcol[f]);
    }
  }
  return 0;
}

int structure(int64_t n, int64_t ncells, int npc, const int32_t* cells, std::vector<int32_t>& rowptr,
              std::vector<int32_t>& col) {
  // node -> cells
  std::vector<int32_t> cstart(n + 1, 0), clist(ncells * npc);
  for (int64_t k = 0; k < ncells * npc; ++k) {
    if (cells[k] < 0 || cells[k] >= n) return fail("cell node out of range");
    cstart[cells[k] + 1]++;
  }
  for (int64_t i = 0; i < n; ++i) cstart[i + 1] += cstart[i];
  {
    std::vector<int32_t> fill(cstart.begin(), cstart.end() - 1);
    for (int64_t c = 0; c < ncells; ++c)
      for (int m = 0; m < npc; ++m) clist[fill[cells[c * npc + m]]++] = (int32_t)c;
  }
  std::vector<int32_t> cnt(n);
  std::vector<std::vector<int32_t>> rows(n);
#pragma omp parallel for schedule(dynamic, 4096)
  for (int64_t i = 0; i < n; ++i) {
    std::vector<int32_t>& r = rows[i];
    for (int32_t k = cstart[i]; k < cstart[i + 1]; ++k) {
      const int32_t* v = cells + (int64_t)clist[k] * npc;
      r.insert(r.end(), v, v + npc);
    }
    std::sort(r.begin(), r.end());
    r.erase(std::unique(r.begin(), r.end()), r.end());
    if (r.empty()) r.push_back((int32_t)i);  // isolated node: diagonal only
    cnt[i] = (int32_t)r.size();
  }
  rowptr.assign(n + 1, 0);
  for (int64_t i = 0; i < n; ++i) rowptr[i + 1] = rowptr[i] + cnt[i];
  col.resize(rowptr[n]);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    std::copy(rows[i].begin(), rows[i].end(), col.begin() + rowptr[i]);
    std::vector<int32_t>().swap(rows[i]);
  }
  return 0;
}

// 1D Q1 element: stiffness and mass on [0,h]
inline double s1(int p, int q, double h) { return (p == q ? 1.0 : -1.0) / h; }
inline double m1(int p, int q, double h) { return h * (p == q ? 2.0 : 1.0) / 6.0; }

int assemble(int dim, int64_t n, const double* X, int64_t ncells, const int32_t* cells,
             const double* kcell, const int32_t* rowptr, const int32_t* col, int64_t ndir,
             const int32_t* dir, double* val) {
  const int npc = dim == 3 ? 8 : 4;
  // per-cell box sizes
  std::vector<double> hx(ncells), hy(ncells), hz(ncells, 1.0);
  for (int64_t c = 0; c < ncells; ++c) {
    const int32_t* v = cells + c * npc;
    const double* x0 = X + 3 * (int64_t)v[0];
    hx[c] = X[3 * (int64_t)v[1]] - x0[0];
    hy[c] = X[3 * (int64_t)v[3] + 1] - x0[1];
    if (dim == 3) hz[c] = X[3 * (int64_t)v[4] + 2] - x0[2];
    // every corner must sit at its box position
    for (int m = 0; m < npc; ++m) {
      const double* xm = X + 3 * (int64_t)v[m];
      const int* p = dim == 3 ? kHexCorner[m] : kQuadCorner[m];
      const double e[3] = {x0[0] + p[0] * hx[c], x0[1] + p[1] * hy[c],
                           dim == 3 ? x0[2] + p[2] * hz[c] : x0[2]};
      for (int d = 0; d < dim; ++d)
        if (std::fabs(xm[d] - e[d]) > 1e-12 * (std::fabs(e[d]) + 1.0))
          return fail("cell %lld is not an axis-aligned box", (long long)c);
    }
    if (hx[c] <= 0 || hy[c] <= 0 || hz[c] <= 0)
      return fail("cell %lld has inverted orientation", (long long)c);
  }
  // node -> (cell, local index) in ascending cell order
  std::vector<int32_t> cstart(n + 1, 0), clist(ncells * npc);
  std::vector<int8_t> cloc(ncells * npc);
  for (int64_t k = 0; k < ncells * npc; ++k) cstart[cells[k] + 1]++;
  for (int64_t i = 0; i < n; ++i) cstart[i + 1] += cstart[i];
  {
    std::vector<int32_t> fill(cstart.begin(), cstart.end() - 1);
    for (int64_t c = 0; c < ncells; ++c)
      for (int m = 0; m < npc; ++m) {
        const int32_t k = fill[cells[c * npc + m]]++;
        clist[k] = (int32_t)c;
        cloc[k] = (int8_t)m;
      }
  }
  std::vector<uint8_t> isdir(n, 0);
  for (int64_t k = 0; k < ndir; ++k) isdir[dir[k]] = 1;
  int bad = 0;
#pragma omp parallel for schedule(dynamic, 4096) reduction(+ : bad)
  for (int64_t i = 0; i < n; ++i) {
    const int32_t rs = rowptr[i], re = rowptr[i + 1];
    for (int32_t p = rs; p < re; ++p) val[p] = 0.0;
    for (int32_t k = cstart[i]; k < cstart[i + 1]; ++k) {
      const int64_t c = clist[k];
      const int a = cloc[k];
      const double kc = kcell ? kcell[c] : 1.0;
      const int32_t* v = cells + c * npc;
      const int* pa = dim == 3 ? kHexCorner[a] : kQuadCorner[a];
      for (int b = 0; b < npc; ++b) {
        const int* pb = dim == 3 ? kHexCorner[b] : kQuadCorner[b];
        double e;
        if (dim == 3)
          e = s1(pa[0], pb[0], hx[c]) * m1(pa[1], pb[1], hy[c]) * m1(pa[2], pb[2], hz[c]) +
              m1(pa[0], pb[0], hx[c]) * s1(pa[1], pb[1], hy[c]) * m1(pa[2], pb[2], hz[c]) +
              m1(pa[0], pb[0], hx[c]) * m1(pa[1], pb[1], hy[c]) * s1(pa[2], pb[2], hz[c]);
        else
          e = s1(pa[0], pb[0], hx[c]) * m1(pa[1], pb[1], hy[c]) +
              m1(pa[0], pb[0], hx[c]) * s1(pa[1], pb[1], hy[c]);
        const int32_t j = v[b];
        const int32_t* it = std::lower_bound(col + rs, col + re, j);
        if (it == col + re || *it != j) {
          bad++;
          continue;
        }
        val[it - col] += kc * e;
      }
    }
    // strong Dirichlet (SparseBlockMatrix::dirichlet): row and column zero, diagonal 1
    for (int32_t p = rs; p < re; ++p) {
      if (isdir[i])
        val[p] = (col[p] == i) ? 1.0 : 0.0;
      else if (isdir[col[p]])
        val[p] = 0.0;
    }
  }
  if (bad) return fail("matrix structure misses %d couplings", bad);
  return 0;
}

}  // namespace

struct ghs_hier {
  std::vector<Mesh> levels;  // 0 = coarsest kept level
};

extern "C" {

const char* ghs_last_error(void) { return g_err.c_str(); }

int ghs_build(const char* inp_path, int dim, int prerefine, int drop_coarse, ghs_hier** out) {
  if (!out || !inp_path || (dim != 2 && dim != 3) || prerefine < 0 || drop_coarse < 0 ||
      drop_coarse > prerefine)
    return fail("ghs_build: bad arguments");
  *out = nullptr;
  std::vector<Mesh> all(1);
  if (read_inp(inp_path, dim, all[0])) return -1;
  for (int r = 0; r < prerefine; ++r) {
    all.emplace_back();
    if (refine(all[r], all[r + 1])) return -1;
  }
  ghs_hier* h = new ghs_hier();
  h->levels.assign(std::make_move_iterator(all.begin() + drop_coarse),
                   std::make_move_iterator(all.end()));
  *out = h;
  return 0;
}

int ghs_destroy(ghs_hier* h) {
  delete h;
  return 0;
}

int ghs_nlevels(const ghs_hier* h) { return h ? (int)h->levels.size() : -1; }

static const Mesh* lev(const ghs_hier* h, int l) {
  if (!h || l < 0 || l >= (int)h->levels.size()) {
    fail("level %d out of range", l);
    return nullptr;
  }
  return &h->levels[l];
}

int ghs_level_sizes(const ghs_hier* h, int l, int64_t* nnodes, int64_t* ncells, int64_t* nnz) {
  const Mesh* M = lev(h, l);
  if (!M) return -1;
  if (nnodes) *nnodes = M->n();
  if (ncells) *ncells = M->ncells();
  if (nnz) {
    std::vector<int32_t> rp, c;
    if (structure(M->n(), M->ncells(), M->npc, M->cells.data(), rp, c)) return -1;
    *nnz = rp.back();
  }
  return 0;
}

int ghs_level_coords(const ghs_hier* h, int l, double* xyz) {
  const Mesh* M = lev(h, l);
  if (!M) return -1;
  std::memcpy(xyz, M->X.data(), M->X.size() * sizeof(double));
  return 0;
}

int ghs_level_cells(const ghs_hier* h, int l, int32_t* cells) {
  const Mesh* M = lev(h, l);
  if (!M) return -1;
  std::memcpy(cells, M->cells.data(), M->cells.size() * sizeof(int32_t));
  return 0;
}

int ghs_level_structure(const ghs_hier* h, int l, int32_t* rowptr, int32_t* col) {
  const Mesh* M = lev(h, l);
  if (!M) return -1;
  std::vector<int32_t> rp, c;
  if (structure(M->n(), M->ncells(), M->npc, M->cells.data(), rp, c)) return -1;
  std::memcpy(rowptr, rp.data(), rp.size() * sizeof(int32_t));
  if (col) std::memcpy(col, c.data(), c.size() * sizeof(int32_t));
  return 0;
}

static void dir_nodes(const Mesh* M, int ncol, const int32_t* cols, std::vector<int32_t>& out) {
  const int per = M->dim == 3 ? 4 : 2;
  std::vector<uint8_t> mark(M->n(), 0);
  for (size_t f = 0; f < M->bcol.size(); ++f) {
    bool sel = false;
    for (int k = 0; k < ncol; ++k) sel |= (M->bcol[f] == cols[k]);
    if (!sel) continue;
    for (int t = 0; t < per; ++t) mark[M->bnd[per * f + t]] = 1;
  }
  out.clear();
  for (int64_t i = 0; i < M->n(); ++i)
    if (mark[i]) out.push_back((int32_t)i);
}

int ghs_level_dirichlet_count(const ghs_hier* h, int l, int ncolours, const int32_t* colours,
                              int64_t* count) {
  const Mesh* M = lev(h, l);
  if (!M) return -1;
  std::vector<int32_t> d;
  dir_nodes(M, ncolours, colours, d);
  *count = (int64_t)d.size();
  return 0;
}

int ghs_level_dirichlet(const ghs_hier* h, int l, int ncolours, const int32_t* colours,
                        int32_t* nodes) {
  const Mesh* M = lev(h, l);
  if (!M) return -1;
  std::vector<int32_t> d;
  dir_nodes(M, ncolours, colours, d);
  std::memcpy(nodes, d.data(), d.size() * sizeof(int32_t));
  return 0;
}

int ghs_level_assemble(const ghs_hier* h, int l, const double* kcell, int ncolours,
                       const int32_t* colours, double* val) {
  const Mesh* M = lev(h, l);
  if (!M) return -1;
  std::vector<int32_t> rp, c, d;
  if (structure(M->n(), M->ncells(), M->npc, M->cells.data(), rp, c)) return -1;
  dir_nodes(M, ncolours, colours, d);
  return assemble(M->dim, M->n(), M->X.data(), M->ncells(), M->cells.data(), kcell, rp.data(),
                  c.data(), (int64_t)d.size(), d.data(), val);
}

int ghs_level_transfer_sizes(const ghs_hier* h, int l, int64_t* ncoarse, int64_t* n2,
                             int64_t* n4, int64_t* n8) {
  const Mesh* M = lev(h, l);
  if (!M) return -1;
  if (l == 0) return fail("level 0 has no coarser level");
  *ncoarse = M->ncoarse;
  *n2 = (int64_t)M->zweier.size() / 3;
  *n4 = (int64_t)M->vierer.size() / 5;
  *n8 = (int64_t)M->achter.size() / 9;
  return 0;
}

int ghs_level_transfer(const ghs_hier* h, int l, int32_t* c2f, int32_t* zweier, int32_t* vierer,
                       int32_t* achter) {
  const Mesh* M = lev(h, l);
  if (!M) return -1;
  if (l == 0) return fail("level 0 has no coarser level");
  for (int64_t i = 0; i < M->ncoarse; ++i) c2f[i] = (int32_t)i;
  std::memcpy(zweier, M->zweier.data(), M->zweier.size() * sizeof(int32_t));
  std::memcpy(vierer, M->vierer.data(), M->vierer.size() * sizeof(int32_t));
  std::memcpy(achter, M->achter.data(), M->achter.size() * sizeof(int32_t));
  return 0;
}

int ghs_structure_from_cells(int64_t n, int64_t ncells, int npc, const int32_t* cells,
                             int32_t* rowptr, int32_t* col) {
  std::vector<int32_t> rp, c;
  if (structure(n, ncells, npc, cells, rp, c)) return -1;
  std::memcpy(rowptr, rp.data(), rp.size() * sizeof(int32_t));
  if (col) std::memcpy(col, c.data(), c.size() * sizeof(int32_t));
  return 0;
}

int ghs_assemble_on(int dim, int64_t n, const double* xyz, int64_t ncells, const int32_t* cells,
                    const double* kcell, const int32_t* rowptr, const int32_t* col, int64_t ndir,
                    const int32_t* dir, double* val) {
  return assemble(dim, n, xyz, ncells, cells, kcell, rowptr, col, ndir, dir, val);
}

}  // extern "C"
