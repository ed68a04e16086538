// Host-side structure producer (see include/gh_structure.h): .inp reader,
// nested uniform refinement of quad/hex meshes, mesh-transfer maps, Dirichlet
// lists, CSR structure, Q1 assembly for -div(k grad u), patches and the
// patch-based Stokes-LPS block assembly.
#include "gh_structure.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
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
  // adaptive patch meshes: refinement generation of each patch (empty:
  // uniform) and the Q2 hanging-node constraints (4 / 10 entries each)
  std::vector<int8_t> plev;
  std::vector<int32_t> hn_e, hn_f;
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

// One refinement (nested numbering): every cell, or the cells with mark[c]
// set (the others are copied; cell order is kept, a refined cell's 8 (4)
// children take its place, so a mark that is uniform over each group of
// sibling cells keeps the patch grouping).
int refine(const Mesh& C, Mesh& F, const std::vector<uint8_t>* mark = nullptr) {
  const int dim = C.dim;
  F.dim = dim;
  F.npc = C.npc;
  const int64_t n0 = C.n(), nc = C.ncells();
  F.X = C.X;
  F.ncoarse = n0;
  std::vector<double>& X = F.X;
  auto marked = [&](int64_t c) { return !mark || (*mark)[c] != 0; };
  std::vector<int32_t> cmid(nc, -1);
  // cell midpoints
  for (int64_t c = 0; c < nc; ++c) {
    if (!marked(c)) continue;
    const int32_t* v = &C.cells[c * C.npc];
    double s[3] = {0, 0, 0};
    for (int m = 0; m < C.npc; ++m)
      for (int d = 0; d < 3; ++d) s[d] += C.X[3 * v[m] + d];
    const int32_t id = (int32_t)(X.size() / 3);
    cmid[c] = id;
    for (int d = 0; d < 3; ++d) X.push_back(s[d] / C.npc);
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
  // With marks, a midpoint may already be a node of C (a hanging node where
  // a neighbour is finer): it is found by its exact coordinates, which the
  // refinement that made it computed the same way, and kept (a coarse node,
  // no 2er/4er entry).
  std::unordered_map<std::string, int32_t> at;
  if (mark) {
    at.reserve(n0 * 2);
    for (int64_t i = 0; i < n0; ++i) at.emplace(std::string((const char*)&C.X[3 * i], 24), (int32_t)i);
  }
  auto existing = [&](const double* x) -> int32_t {
    if (!mark) return -1;
    auto it = at.find(std::string((const char*)x, 24));
    return it == at.end() ? -1 : it->second;
  };
  // face midpoints (3D)
  std::unordered_map<uint64_t, int32_t> fmap, emap;
  fmap.reserve(dim == 3 ? nc * 4 : 1);
  emap.reserve(nc * (dim == 3 ? 4 : 3));
  if (dim == 3) {
    for (int64_t c = 0; c < nc; ++c) {
      if (!marked(c)) continue;
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
          double s[3] = {0, 0, 0};
          for (int t = 0; t < 4; ++t)
            for (int d = 0; d < 3; ++d) s[d] += C.X[3 * q[t] + d];
          for (int d = 0; d < 3; ++d) s[d] /= 4.0;
          const int32_t old = existing(s);
          if (old >= 0) {
            fmap.emplace(key, old);
            continue;
          }
          const int32_t id = (int32_t)(X.size() / 3);
          fmap.emplace(key, id);
          for (int d = 0; d < 3; ++d) X.push_back(s[d]);
          F.vierer.push_back(id);
          for (int t = 0; t < 4; ++t) F.vierer.push_back(q[t]);
        }
    }
  }
  // edge midpoints
  for (int64_t c = 0; c < nc; ++c) {
    if (!marked(c)) continue;
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
      double mid[3];
      for (int d = 0; d < 3; ++d) mid[d] = 0.5 * (C.X[3 * a + d] + C.X[3 * b + d]);
      const int32_t old = existing(mid);
      if (old >= 0) {
        emap.emplace(key, old);
        continue;
      }
      const int32_t id = (int32_t)(X.size() / 3);
      emap.emplace(key, id);
      for (int d = 0; d < 3; ++d) X.push_back(mid[d]);
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
    if (!marked(c)) {
      F.cells.insert(F.cells.end(), v, v + C.npc);
      continue;
    }
    if (dim == 3) {
      int32_t G[3][3][3];
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b)
          for (int k = 0; k < 3; ++k) {
            const int odd = (a & 1) + (b & 1) + (k & 1);
            if (odd == 0) {
              G[a][b][k] = corner(v, a / 2, b / 2, k / 2);
            } else if (odd == 3) {
              G[a][b][k] = cmid[c];
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
            G[a][b] = cmid[c];
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
      if (fit == fmap.end()) {
        if (!mark) return fail("boundary quad %lld is not a cell face", (long long)f);
        F.bnd.insert(F.bnd.end(), q, q + 4);  // face of a cell left as it is
        F.bcol.push_back(C.bcol[f]);
        continue;
      }
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
      if (it == emap.end()) {
        if (!mark) return fail("boundary line %lld is not a cell edge", (long long)f);
        F.bnd.insert(F.bnd.end(), {a, b});
        F.bcol.push_back(C.bcol[f]);
        continue;
      }
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

// Patches of a refined level: the children of coarse cell p are cells
// 8p..8p+7 (4p..4p+3 in 2D), child (i, j, k) = 8p + i + 2j + 4k.  Patch node
// (a, b, c), a, b, c in {0, 1, 2}, is stored at a + 3b + 9c (x fastest), the
// element order of Gascoigne's patch dof handler (the Q1-patch basis
// src/Discretization/Q2/baseq13dpatch.h:58-68,104-140 places child k's corner
// i at perm[i] + perm[k] with perm = 0,1,3,4,9,10,12,13).
int patches(const Mesh& M, std::vector<int32_t>& P) {
  if (M.ncoarse == 0) return fail("the coarsest mesh of a hierarchy has no patches");
  const int dim = M.dim, npc = M.npc;
  const int64_t nch = dim == 3 ? 8 : 4;
  const int nloc = dim == 3 ? 27 : 9;
  const int64_t np = M.ncells() / nch;
  if (np * nch != M.ncells()) return fail("cell count is not a multiple of %d", (int)nch);
  P.assign(np * nloc, -1);
  for (int64_t p = 0; p < np; ++p) {
    int32_t* out = &P[p * nloc];
    for (int64_t ch = 0; ch < nch; ++ch) {
      const int i = ch & 1, j = (ch >> 1) & 1, k = dim == 3 ? (int)(ch >> 2) : 0;
      const int32_t* v = &M.cells[(p * nch + ch) * npc];
      for (int m = 0; m < npc; ++m) {
        const int a = i + (dim == 3 ? kHexCorner[m][0] : kQuadCorner[m][0]);
        const int b = j + (dim == 3 ? kHexCorner[m][1] : kQuadCorner[m][1]);
        const int c = k + (dim == 3 ? kHexCorner[m][2] : 0);
        int32_t& slot = out[a + 3 * b + 9 * c];
        if (slot >= 0 && slot != v[m]) return fail("patch %lld: children do not share nodes", (long long)p);
        slot = v[m];
      }
    }
    for (int a = 0; a < nloc; ++a)
      if (out[a] < 0) return fail("patch %lld: node %d missing", (long long)p, a);
  }
  return 0;
}

// 1D matrices of the piecewise-linear basis on a patch interval [0, 2h] with
// nodes 0, h, 2h: stiffness S, mass M, and C(a, b) = int phi_a phi_b'.
struct Patch1d {
  double S[3][3], M[3][3], C[3][3];
  explicit Patch1d(double h) {
    const double s[3][3] = {{1, -1, 0}, {-1, 2, -1}, {0, -1, 1}};
    const double m[3][3] = {{2, 1, 0}, {1, 4, 1}, {0, 1, 2}};
    const double c[3][3] = {{-0.5, 0.5, 0}, {-0.5, 0, 0.5}, {0, -0.5, 0.5}};
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) {
        S[a][b] = s[a][b] / h;
        M[a][b] = h * m[a][b] / 6.0;
        C[a][b] = c[a][b];
      }
  }
};

// Stokes-LPS patch element of NavierStokesLps<3> at U = 0 with
// symmetrictensor = 0 (src/Problems/navierstokes.h:144-168 Galerkin part,
// :251-268 pressure stabilisation, :221-227 lpspoint) on an axis-aligned box
// patch of 2x2x2 cells of size hx x hy x hz.  Components (p, u, v, w); the
// 4x4 block of row node a (test function) and column node b (trial function):
//   (0, d+1)   +int dphi_b/dx_d phi_a                     divergence
//   (d+1, d+1) +visc int grad phi_b . grad phi_a           viscosity
//   (d+1, 0)   -int phi_b dphi_a/dx_d                      pressure
//   (0, 0)     +alpha_P int grad (I-pi)phi_b . grad (I-pi)phi_a
// with pi the trilinear interpolation from the 8 patch corners (the
// fluctuation operator of ElementLpsIntegrator::Projection,
// src/DofHandler/elementlpsintegrator.h:62-105) and
// alpha_P = alpha / (visc / h_P^2), h_P = (patch volume)^(1/3)
// (Volume2MeshSize of the patch Jacobian, elementintegrator.h:53-56).  The
// integrals are exact (the reference's 2x2x2 Gauss rule per cell is exact for
// these integrands), so they are formed as tensor products of 1D matrices.
struct StokesPatch {
  double blk[27][27][16];
  StokesPatch(double hx, double hy, double hz, double visc, double alpha) {
    const Patch1d X(hx), Y(hy), Z(hz);
    double K[27][27], D[3][27][27];
    for (int a = 0; a < 27; ++a)
      for (int b = 0; b < 27; ++b) {
        const int ax = a % 3, ay = (a / 3) % 3, az = a / 9;
        const int bx = b % 3, by = (b / 3) % 3, bz = b / 9;
        K[a][b] = X.S[ax][bx] * Y.M[ay][by] * Z.M[az][bz] + X.M[ax][bx] * Y.S[ay][by] * Z.M[az][bz] +
                  X.M[ax][bx] * Y.M[ay][by] * Z.S[az][bz];
        D[0][a][b] = X.C[ax][bx] * Y.M[ay][by] * Z.M[az][bz];
        D[1][a][b] = X.M[ax][bx] * Y.C[ay][by] * Z.M[az][bz];
        D[2][a][b] = X.M[ax][bx] * Y.M[ay][by] * Z.C[az][bz];
      }
    // F = I - Pi: column b holds the nodal values of (I - pi) phi_b.  pi phi_b
    // is zero unless b is a corner, where it is the coarse trilinear hat
    // (1 at b, 1/2 on the adjacent edge midpoints, 1/4 on the adjacent face
    // centres, 1/8 at the patch centre).
    double F[27][27];
    const double p1[3][3] = {{1, 0, 0}, {0.5, 0, 0.5}, {0, 0, 1}};  // p1[node][corner]
    for (int a = 0; a < 27; ++a)
      for (int b = 0; b < 27; ++b) {
        const int ax = a % 3, ay = (a / 3) % 3, az = a / 9;
        const int bx = b % 3, by = (b / 3) % 3, bz = b / 9;
        F[a][b] = (a == b ? 1.0 : 0.0) - p1[ax][bx] * p1[ay][by] * p1[az][bz];
      }
    double KF[27][27];
    for (int a = 0; a < 27; ++a)
      for (int b = 0; b < 27; ++b) {
        double s = 0;
        for (int k = 0; k < 27; ++k) s += K[a][k] * F[k][b];
        KF[a][b] = s;
      }
    const double hP = std::cbrt(8.0 * hx * hy * hz);
    const double alphaP = alpha / (visc / hP / hP);
    for (int a = 0; a < 27; ++a)
      for (int b = 0; b < 27; ++b) {
        double L = 0;
        for (int k = 0; k < 27; ++k) L += F[k][a] * KF[k][b];
        double* e = blk[a][b];
        for (int t = 0; t < 16; ++t) e[t] = 0.0;
        e[0] = alphaP * L;
        for (int d = 0; d < 3; ++d) {
          e[0 * 4 + d + 1] = D[d][a][b];
          e[(d + 1) * 4 + d + 1] = visc * K[a][b];
          e[(d + 1) * 4 + 0] = -D[d][b][a];
        }
      }
  }
};

// Box sizes of a patch from its lexicographic nodes; fails unless the patch
// is an axis-aligned box split into 8 equal cells.
int patch_box(const Mesh& M, const int32_t* P, double h[3]) {
  const double* x0 = &M.X[3 * (int64_t)P[0]];
  h[0] = M.X[3 * (int64_t)P[1]] - x0[0];
  h[1] = M.X[3 * (int64_t)P[3] + 1] - x0[1];
  h[2] = M.X[3 * (int64_t)P[9] + 2] - x0[2];
  for (int a = 0; a < 27; ++a) {
    const int idx[3] = {a % 3, (a / 3) % 3, a / 9};
    for (int d = 0; d < 3; ++d) {
      const double e = x0[d] + idx[d] * h[d];
      if (std::fabs(M.X[3 * (int64_t)P[a] + d] - e) > 1e-12 * (std::fabs(e) + 1.0)) return -1;
    }
  }
  return (h[0] > 0 && h[1] > 0 && h[2] > 0) ? 0 : -1;
}

// Laplace on the triquadratic patch elements of M, unconstrained, on the patch
// structure (rp, c) of M.
int q2_matrix(const Mesh* M, std::vector<int32_t>& rp, std::vector<int32_t>& c,
              std::vector<double>& vals) {
  if (M->dim != 3) return fail("Q2 patch assembly is three-dimensional");
  std::vector<int32_t> P;
  if (patches(*M, P)) return -1;
  const int64_t n = M->n(), np = (int64_t)P.size() / 27;
  if (structure(n, np, 27, P.data(), rp, c)) return -1;
  // Laplace<3> on the triquadratic patch element (CGDiscQ23d: the 27 patch
  // nodes carry the Lagrange basis of BaseQ23d; src/Problems/laplace.h,
  // src/DofHandler/cgdisc.h:1291): per direction the quadratic element on
  // [0, L], L = 2h, nodes 0, L/2, L, with stiffness (7,-8,1;-8,16,-8;1,-8,7)/(3L)
  // and mass (4,2,-1;2,16,2;-1,2,4) L/30, exact (the reference's 3x3x3 Gauss
  // rule is exact for these products)
  auto q2 = [](double L, double S[3][3], double Mm[3][3]) {
    const double s[3][3] = {{7, -8, 1}, {-8, 16, -8}, {1, -8, 7}};
    const double m[3][3] = {{4, 2, -1}, {2, 16, 2}, {-1, 2, 4}};
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) {
        S[a][b] = s[a][b] / (3.0 * L);
        Mm[a][b] = m[a][b] * L / 30.0;
      }
  };
  std::vector<int32_t> shape(np);
  std::vector<std::array<double, 3>> sizes;
  std::vector<std::vector<double>> elem;  // 27 x 27 per shape
  for (int64_t p = 0; p < np; ++p) {
    double hb[3];
    if (patch_box(*M, &P[p * 27], hb)) return fail("patch %lld is not an axis-aligned box of 8 equal cells", (long long)p);
    int sidx = 0;
    while (sidx < (int)sizes.size() && !(sizes[sidx][0] == hb[0] && sizes[sidx][1] == hb[1] && sizes[sidx][2] == hb[2])) ++sidx;
    if (sidx == (int)sizes.size()) {
      if (sidx >= 64) return fail("Q2 assembly: more than 64 distinct patch sizes");
      sizes.push_back({hb[0], hb[1], hb[2]});
      double SX[3][3], MX[3][3], SY[3][3], MY[3][3], SZ[3][3], MZ[3][3];
      q2(2 * hb[0], SX, MX);
      q2(2 * hb[1], SY, MY);
      q2(2 * hb[2], SZ, MZ);
      std::vector<double> E(27 * 27);
      for (int a = 0; a < 27; ++a)
        for (int b = 0; b < 27; ++b) {
          const int ax = a % 3, ay = (a / 3) % 3, az = a / 9;
          const int bx = b % 3, by = (b / 3) % 3, bz = b / 9;
          E[a * 27 + b] = SX[ax][bx] * MY[ay][by] * MZ[az][bz] + MX[ax][bx] * SY[ay][by] * MZ[az][bz] +
                          MX[ax][bx] * MY[ay][by] * SZ[az][bz];
        }
      elem.push_back(std::move(E));
    }
    shape[p] = sidx;
  }
  std::vector<int32_t> pstart(n + 1, 0), plist(np * 27);
  std::vector<int8_t> ploc(np * 27);
  for (int64_t k = 0; k < np * 27; ++k) pstart[P[k] + 1]++;
  for (int64_t i = 0; i < n; ++i) pstart[i + 1] += pstart[i];
  {
    std::vector<int32_t> fill(pstart.begin(), pstart.end() - 1);
    for (int64_t p = 0; p < np; ++p)
      for (int a = 0; a < 27; ++a) {
        const int32_t k = fill[P[p * 27 + a]]++;
        plist[k] = (int32_t)p;
        ploc[k] = (int8_t)a;
      }
  }
  vals.assign(rp[n], 0.0);
  double* val = vals.data();
  int bad = 0;
#pragma omp parallel for schedule(dynamic, 1024) reduction(+ : bad)
  for (int64_t i = 0; i < n; ++i) {
    const int32_t rs = rp[i], re = rp[i + 1];
    for (int32_t q = rs; q < re; ++q) val[q] = 0.0;
    for (int32_t k = pstart[i]; k < pstart[i + 1]; ++k) {
      const int64_t p = plist[k];
      const int a = ploc[k];
      const double* E = elem[shape[p]].data() + a * 27;
      for (int b = 0; b < 27; ++b) {
        const int32_t j = P[p * 27 + b];
        const int32_t* it = std::lower_bound(c.data() + rs, c.data() + re, j);
        if (it == c.data() + re || *it != j) {
          bad++;
          continue;
        }
        val[it - c.data()] += E[b];
      }
    }
  }
  if (bad) return fail("patch structure misses %d couplings", bad);
  return 0;
}

// ------------------------------------------------ adaptive Q2 patch meshes
// Integer lattice positions of the nodes (coordinates must be multiples of
// 2^-20 in [0, 2): the refined unit-box meshes of the benchmark workloads).
int lattice(const Mesh& M, std::vector<std::array<int32_t, 3>>& L) {
  const double s = 1048576.0;
  L.resize(M.n());
  for (int64_t i = 0; i < M.n(); ++i)
    for (int d = 0; d < 3; ++d) {
      const double v = M.X[3 * i + d] * s;
      const double r = std::nearbyint(v);
      if (std::fabs(v - r) > 1e-6 || r < 0 || r >= 2097152.0)
        return fail("adaptive Q2 meshes need node coordinates on the dyadic lattice of [0,2)^3");
      L[i][d] = (int32_t)r;
    }
  return 0;
}
inline uint64_t lkey(int64_t a, int64_t b, int64_t c) {
  return ((uint64_t)a << 42) | ((uint64_t)b << 21) | (uint64_t)c;
}

// Q2 hanging nodes of a patch mesh: a node at a half position of a patch's
// face (between two of its Q2 nodes) is interpolated from that patch, as in
// HNStructureQ23d (src/Discretization/Q2/hnstructureq23d.cc:120-156):
//   edge constraint [h, near, mid, far] on a Q2 line of the face, weights
//     3/8, 3/4, -1/8 (the quadratic Lagrange basis at a quarter point);
//   face constraint [h, g_0..g_8], g_{3a+b} the Q2 node at position a
//     (near, mid, far) along one face direction and b along the other,
//     weight w_a w_b.
int q2_hanging(const Mesh& M, std::vector<int32_t>& E, std::vector<int32_t>& Fc) {
  E.clear();
  Fc.clear();
  std::vector<int32_t> P;
  if (patches(M, P)) return -1;
  std::vector<std::array<int32_t, 3>> L;
  if (lattice(M, L)) return -1;
  std::unordered_map<uint64_t, int32_t> at;
  at.reserve(M.n() * 2);
  for (int64_t i = 0; i < M.n(); ++i) at.emplace(lkey(L[i][0], L[i][1], L[i][2]), (int32_t)i);
  std::vector<int8_t> done(M.n(), 0);
  const int64_t np = (int64_t)P.size() / 27;
  for (int64_t p = 0; p < np; ++p) {
    const int32_t* pn = &P[p * 27];
    const std::array<int32_t, 3>& o = L[pn[0]];
    const int32_t cs = L[pn[1]][0] - o[0];  // Q2 node spacing (a cell) in lattice units
    if (cs <= 0 || (cs & 1) || L[pn[3]][1] - o[1] != cs || L[pn[9]][2] - o[2] != cs)
      return fail("patch %lld is not a cube of 8 equal cells on the lattice", (long long)p);
    auto node = [&](int i0, int i1, int i2) { return pn[i0 + 3 * i1 + 9 * i2]; };
    for (int ax = 0; ax < 3; ++ax)
      for (int side = 0; side <= 2; side += 2) {
        const int u = (ax + 1) % 3, w = (ax + 2) % 3;
        for (int iu = 0; iu <= 4; ++iu)
          for (int iw = 0; iw <= 4; ++iw) {
            if (!(iu & 1) && !(iw & 1)) continue;  // a Q2 node of the patch
            int64_t q[3];
            q[ax] = o[ax] + (int64_t)side * cs;
            q[u] = o[u] + (int64_t)iu * (cs / 2);
            q[w] = o[w] + (int64_t)iw * (cs / 2);
            auto it = at.find(lkey(q[0], q[1], q[2]));
            if (it == at.end()) continue;
            const int32_t h = it->second;
            if (done[h]) continue;
            done[h] = 1;
            // per in-face direction: near end, middle, far end of the Q2 line
            auto line = [](int i, int (&idx)[3]) {
              const int nr = (i == 1) ? 0 : 2;
              idx[0] = nr;
              idx[1] = 1;
              idx[2] = 2 - nr;
            };
            int I[3];
            I[ax] = side;
            if ((iu & 1) && (iw & 1)) {
              int U[3], W[3];
              line(iu, U);
              line(iw, W);
              Fc.push_back(h);
              for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) {
                  I[u] = U[a];
                  I[w] = W[b];
                  Fc.push_back(node(I[0], I[1], I[2]));
                }
            } else {
              const int dir = (iu & 1) ? u : w;
              const int other = (iu & 1) ? w : u;
              int D[3];
              line((iu & 1) ? iu : iw, D);
              I[other] = ((iu & 1) ? iw : iu) / 2;
              E.push_back(h);
              for (int a = 0; a < 3; ++a) {
                I[dir] = D[a];
                E.push_back(node(I[0], I[1], I[2]));
              }
            }
          }
      }
  }
  // no constraint may depend on another hanging node
  for (size_t k = 0; k < E.size(); k += 4)
    for (int t = 1; t < 4; ++t)
      if (done[E[k + t]]) return fail("hanging node %d has a hanging parent", E[k]);
  for (size_t k = 0; k < Fc.size(); k += 10)
    for (int t = 1; t < 10; ++t)
      if (done[Fc[k + t]]) return fail("hanging node %d has a hanging parent", Fc[k]);
  return 0;
}

// One adaptive step on a patch mesh, marking as MeshAgent::refine_nodes does
// for the nodes inside the ball (|x - c| < r) with patch depth 1 and the
// default etapatcher (HierarchicalMesh::vertex_patch_refine,
// src/Mesh/hierarchicalmesh.cc:272-290; HierarchicalMesh3d::FillVertexLevels,
// RefineCoarseNodes, VertexToCells, LoadFathers3d, patch_refine,
// regular_grid3d_three_refine, src/Mesh/hierarchicalmesh3d.cc:1938-1998,
// 2081-2254; CoarseHierarchicalMesh3d::refine,
// src/Mesh/coarsehierarchicalmesh3d.cc:94-130).  The cells of one patch are
// the children of one father and share its generation (level), so on the
// patch mesh the steps read:
//   1. vertex level: the lowest level of the cells at a node;
//   2. for every cell with a marked node, its patch's corners of the lowest
//      vertex level become coarse marks;
//   3. a cell with a coarse-marked corner of its own level marks its patch;
//   4. on the mesh of patches (the coarsened mesh, corners as vertices) the
//      closure: wherever the levels at a corner would differ by 2 or more
//      after the marked patches are refined, the patches there more than one
//      level below the finest are marked too, until nothing changes;
//   5. every cell of a marked patch is refined.
// Refining the children of the marked patches on the full mesh then needs no
// further closure: at any corner the patch levels differ by at most one.
int adapt(const Mesh& C, Mesh& F, double r, const double ctr[3]) {
  std::vector<int32_t> P;
  if (patches(C, P)) return -1;
  const int64_t np = (int64_t)P.size() / 27;
  const int64_t nc = C.ncells();
  std::vector<int8_t> lev = C.plev;
  if ((int64_t)lev.size() != np) lev.assign(np, 0);
  const int corners[8] = {0, 2, 6, 8, 18, 20, 24, 26};
  // 1. vertex levels
  std::vector<int32_t> vlev(C.n(), 1 << 20);
  for (int64_t c = 0; c < nc; ++c)
    for (int m = 0; m < 8; ++m) {
      const int32_t v = C.cells[c * 8 + m];
      vlev[v] = std::min<int32_t>(vlev[v], lev[c / 8]);
    }
  std::vector<uint8_t> refnode(C.n(), 0);
  for (int64_t i = 0; i < C.n(); ++i) {
    const double* x = &C.X[3 * i];
    const double d2 = (x[0] - ctr[0]) * (x[0] - ctr[0]) + (x[1] - ctr[1]) * (x[1] - ctr[1]) +
                      (x[2] - ctr[2]) * (x[2] - ctr[2]);
    refnode[i] = d2 < r * r;
  }
  // 2. coarse marks: the lowest-level corners of the patch of a marked cell
  std::vector<uint8_t> coarse(C.n(), 0);
  for (int64_t c = 0; c < nc; ++c) {
    bool any = false;
    for (int m = 0; m < 8 && !any; ++m) any = refnode[C.cells[c * 8 + m]];
    if (!any) continue;
    const int32_t* pc = &P[(c / 8) * 27];
    int32_t mn = vlev[pc[corners[0]]];
    for (int t = 1; t < 8; ++t) mn = std::min(mn, vlev[pc[corners[t]]]);
    for (int t = 0; t < 8; ++t)
      if (vlev[pc[corners[t]]] == mn) coarse[pc[corners[t]]] = 1;
  }
  // 3. cells with a coarse-marked corner of their own level mark their patch
  std::vector<uint8_t> pmark(np, 0);
  for (int64_t c = 0; c < nc; ++c)
    for (int m = 0; m < 8; ++m) {
      const int32_t v = C.cells[c * 8 + m];
      if (coarse[v] && vlev[v] == lev[c / 8]) pmark[c / 8] = 1;
    }
  // 4. closure on the mesh of patches
  std::vector<int32_t> vmax(C.n()), vmin(C.n());
  for (;;) {
    std::fill(vmax.begin(), vmax.end(), -1);
    std::fill(vmin.begin(), vmin.end(), 1 << 20);
    for (int64_t p = 0; p < np; ++p) {
      const int32_t l = lev[p] + (pmark[p] ? 1 : 0);
      for (int t = 0; t < 8; ++t) {
        const int32_t v = P[p * 27 + corners[t]];
        vmax[v] = std::max(vmax[v], l);
        vmin[v] = std::min(vmin[v], l);
      }
    }
    int64_t added = 0;
    for (int64_t p = 0; p < np; ++p) {
      if (pmark[p]) continue;
      for (int t = 0; t < 8; ++t) {
        const int32_t v = P[p * 27 + corners[t]];
        if (vmax[v] >= vmin[v] + 2 && vmax[v] >= lev[p] + 2) {
          pmark[p] = 1;
          ++added;
          break;
        }
      }
    }
    if (!added) break;
  }
  // 5. refine the cells of the marked patches
  std::vector<uint8_t> cmark(nc, 0);
  for (int64_t c = 0; c < nc; ++c) cmark[c] = pmark[c / 8];
  if (refine(C, F, &cmark)) return -1;
  F.plev.clear();
  for (int64_t p = 0; p < np; ++p) {
    if (pmark[p])
      F.plev.insert(F.plev.end(), 8, (int8_t)(lev[p] + 1));
    else
      F.plev.push_back(lev[p]);
  }
  return q2_hanging(F, F.hn_e, F.hn_f);
}

// Condensed Q2 operator of a patch mesh with hanging nodes: A_c = T^T A T,
// T the interpolation that sets hanging values from their parents (the
// element-wise condensation of CGDisc::LocalToGlobal with CondenseHanging,
// src/DofHandler/cgdisc.h:353-366, in global form); hanging rows and columns
// are then identity rows (as in the reference's assembled matrices), and the
// Dirichlet rows and columns of the given nodes likewise.
int condense(int64_t n, const std::vector<int32_t>& rp, const std::vector<int32_t>& cl,
             const std::vector<double>& vl, const std::vector<int32_t>& E,
             const std::vector<int32_t>& Fc, const std::vector<uint8_t>& isdir,
             std::vector<int32_t>& rpo, std::vector<int32_t>& clo, std::vector<double>& vlo) {
  // hanging node -> (parent, weight)
  std::vector<int32_t> hstart(n + 1, 0);
  std::vector<std::pair<int32_t, double>> hpar;
  std::vector<std::vector<std::pair<int32_t, double>>> tmp(n);
  const double w1[3] = {0.375, 0.75, -0.125};
  for (size_t k = 0; k < E.size(); k += 4)
    for (int t = 0; t < 3; ++t) tmp[E[k]].push_back({E[k + 1 + t], w1[t]});
  for (size_t k = 0; k < Fc.size(); k += 10)
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) tmp[Fc[k]].push_back({Fc[k + 1 + 3 * a + b], w1[a] * w1[b]});
  for (int64_t i = 0; i < n; ++i) {
    hstart[i + 1] = hstart[i] + (int32_t)tmp[i].size();
    hpar.insert(hpar.end(), tmp[i].begin(), tmp[i].end());
  }
  std::vector<std::vector<std::pair<int32_t, double>>>().swap(tmp);
  auto hanging = [&](int64_t i) { return hstart[i + 1] > hstart[i]; };
  // parent -> hanging rows that feed it (for the row condensation)
  std::vector<int32_t> fstart(n + 1, 0), flist;
  std::vector<double> fw;
  for (const auto& e : hpar) fstart[e.first + 1]++;
  for (int64_t i = 0; i < n; ++i) fstart[i + 1] += fstart[i];
  flist.resize(hpar.size());
  fw.resize(hpar.size());
  {
    std::vector<int32_t> fill(fstart.begin(), fstart.end() - 1);
    for (int64_t h = 0; h < n; ++h)
      for (int32_t k = hstart[h]; k < hstart[h + 1]; ++k) {
        const int32_t q = fill[hpar[k].first]++;
        flist[q] = (int32_t)h;
        fw[q] = hpar[k].second;
      }
  }
  // B = A T, row by row (hanging columns replaced by their parents)
  auto brow = [&](int64_t i, double s, std::vector<std::pair<int32_t, double>>& out) {
    for (int32_t q = rp[i]; q < rp[i + 1]; ++q) {
      const int32_t j = cl[q];
      if (hanging(j)) {
        for (int32_t k = hstart[j]; k < hstart[j + 1]; ++k)
          out.push_back({hpar[k].first, s * (vl[q] * hpar[k].second)});
      } else {
        out.push_back({j, s * vl[q]});
      }
    }
  };
  std::vector<std::vector<std::pair<int32_t, double>>> rows(n);
#pragma omp parallel for schedule(dynamic, 512)
  for (int64_t i = 0; i < n; ++i) {
    std::vector<std::pair<int32_t, double>>& r = rows[i];
    if (hanging(i)) {
      r.push_back({(int32_t)i, 1.0});
      continue;
    }
    brow(i, 1.0, r);
    for (int32_t k = fstart[i]; k < fstart[i + 1]; ++k) brow(flist[k], fw[k], r);
    std::stable_sort(r.begin(), r.end(),
                     [](const std::pair<int32_t, double>& a, const std::pair<int32_t, double>& b) {
                       return a.first < b.first;
                     });
    size_t m = 0;
    for (size_t t = 0; t < r.size(); ++t) {
      if (m && r[m - 1].first == r[t].first)
        r[m - 1].second += r[t].second;
      else
        r[m++] = r[t];
    }
    r.resize(m);
    if (isdir[i]) {
      for (auto& e : r) e.second = (e.first == i) ? 1.0 : 0.0;
    } else {
      for (auto& e : r)
        if (isdir[e.first]) e.second = 0.0;
    }
  }
  rpo.assign(n + 1, 0);
  for (int64_t i = 0; i < n; ++i) rpo[i + 1] = rpo[i] + (int32_t)rows[i].size();
  clo.resize(rpo[n]);
  vlo.resize(rpo[n]);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    int32_t q = rpo[i];
    for (const auto& e : rows[i]) {
      clo[q] = e.first;
      vlo[q++] = e.second;
    }
    std::vector<std::pair<int32_t, double>>().swap(rows[i]);
  }
  return 0;
}

}  // namespace

struct ghs_hier {
  std::vector<Mesh> levels;  // 0 = coarsest kept level
  // condensed Q2 operators between the size query and the copy-out
  struct Condensed {
    std::vector<int32_t> rp, col;
    std::vector<double> val;
  };
  std::map<int, Condensed> q2c;
};

// ------------------------------------ multigrid levels of an adaptive mesh
// A hexahedral cell of the refined unit box is (level L, i, j, k): the cube
// [i, i+1] x [j, j+1] x [k, k+1] * 2^-L.  Packed: 5 bits level, 19 bits each.
inline uint64_t ckey(int L, int64_t i, int64_t j, int64_t k) {
  return ((uint64_t)L << 57) | ((uint64_t)i << 38) | ((uint64_t)j << 19) | (uint64_t)k;
}
inline int ck_level(uint64_t c) { return (int)(c >> 57); }
inline int64_t ck_i(uint64_t c, int d) { return (int64_t)((c >> (38 - 19 * d)) & 0x7FFFF); }

int cell_keys(const Mesh& M, std::vector<uint64_t>& K) {
  std::vector<std::array<int32_t, 3>> L;
  if (lattice(M, L)) return -1;
  K.resize(M.ncells());
  for (int64_t c = 0; c < M.ncells(); ++c) {
    const int32_t* v = &M.cells[c * 8];
    const std::array<int32_t, 3>& o = L[v[0]];
    const int32_t sz = L[v[1]][0] - o[0];
    int lev = 0;
    while (lev <= 20 && (1 << (20 - lev)) != sz) ++lev;
    if (lev > 19 || o[0] % sz || o[1] % sz || o[2] % sz || L[v[6]][1] - o[1] != sz ||
        L[v[6]][2] - o[2] != sz)
      return fail("adaptive levels need cubic cells of the dyadic unit box (cell %lld)",
                  (long long)c);
    K[c] = ckey(lev, o[0] / sz, o[1] / sz, o[2] / sz);
  }
  return 0;
}

// The next coarser multigrid level of a patch mesh given as its cells, as
// GascoigneMeshConstructor builds it (LevelMesh3d::construct_lists,
// src/Mesh/levelmesh3d.cc:703-831, with LevelJumper, src/Mesh/leveljumper.h):
// a grandfather whose 64 grandchildren are all cells of the level is
// represented by its 8 children, every other cell is kept; then, while some
// corner carries cells whose levels differ by more than one, the coarsened
// groups with a corner where a cell is more than one level finer are undone.
void coarser_level(const std::unordered_map<uint64_t, int>& S, std::vector<uint64_t>& out) {
  std::unordered_map<uint64_t, int> opis, oldc, newc;
  for (const auto& kv : S) {
    const uint64_t c = kv.first;
    const int L = ck_level(c);
    if (L < 2) {
      oldc[c] = 1;
      continue;
    }
    opis[ckey(L - 2, ck_i(c, 0) >> 2, ck_i(c, 1) >> 2, ck_i(c, 2) >> 2)] = 1;
  }
  auto grandchildren = [&](uint64_t g, bool present_only, std::vector<uint64_t>& v) {
    v.clear();
    const int L = ck_level(g) + 2;
    for (int a = 0; a < 4; ++a)
      for (int b = 0; b < 4; ++b)
        for (int e = 0; e < 4; ++e) {
          const uint64_t c = ckey(L, 4 * ck_i(g, 0) + a, 4 * ck_i(g, 1) + b, 4 * ck_i(g, 2) + e);
          if (!present_only || S.count(c)) v.push_back(c);
        }
  };
  std::vector<uint64_t> gc;
  for (const auto& kv : opis) {
    const uint64_t g = kv.first;
    grandchildren(g, false, gc);
    bool all = true;
    for (uint64_t c : gc) all = all && S.count(c);
    if (all) {
      const int L = ck_level(g) + 1;
      for (int a = 0; a < 8; ++a)
        newc[ckey(L, 2 * ck_i(g, 0) + (a & 1), 2 * ck_i(g, 1) + ((a >> 1) & 1),
                  2 * ck_i(g, 2) + (a >> 2))] = 1;
    } else {
      grandchildren(g, true, gc);
      for (uint64_t c : gc) oldc[c] = 1;
    }
  }
  // corners in units of the finest lattice (level 19)
  auto corner = [](uint64_t c, int m) {
    const int s = 19 - ck_level(c);
    return ckey(0, (ck_i(c, 0) + kHexCorner[m][0]) << s, (ck_i(c, 1) + kHexCorner[m][1]) << s,
                (ck_i(c, 2) + kHexCorner[m][2]) << s);
  };
  for (;;) {
    std::unordered_map<uint64_t, std::pair<int, int>> phi;  // corner -> (min, max) level
    auto upd = [&](uint64_t c) {
      for (int m = 0; m < 8; ++m) {
        auto it = phi.emplace(corner(c, m), std::make_pair(1000, 0)).first;
        it->second.first = std::min(it->second.first, ck_level(c));
        it->second.second = std::max(it->second.second, ck_level(c));
      }
    };
    for (const auto& kv : newc) upd(kv.first);
    for (const auto& kv : oldc) upd(kv.first);
    bool jump = false;
    for (const auto& kv : phi) jump = jump || kv.second.second > kv.second.first + 1;
    if (!jump) break;
    std::vector<uint64_t> undo;
    for (const auto& kv : newc) {
      const uint64_t q = kv.first;
      for (int m = 0; m < 8; ++m)
        if (phi[corner(q, m)].second > ck_level(q) + 1) {
          undo.push_back(ckey(ck_level(q) - 1, ck_i(q, 0) >> 1, ck_i(q, 1) >> 1, ck_i(q, 2) >> 1));
          break;
        }
    }
    for (uint64_t f : undo) {
      const int L = ck_level(f) + 1;
      for (int a = 0; a < 8; ++a)
        newc.erase(ckey(L, 2 * ck_i(f, 0) + (a & 1), 2 * ck_i(f, 1) + ((a >> 1) & 1),
                        2 * ck_i(f, 2) + (a >> 2)));
      grandchildren(f, true, gc);
      for (uint64_t c : gc) oldc[c] = 1;
    }
  }
  out.clear();
  for (const auto& kv : newc) out.push_back(kv.first);
  for (const auto& kv : oldc) out.push_back(kv.first);
  std::sort(out.begin(), out.end());
}

// The multigrid levels of the adaptive mesh `fine` (all its cells at levels
// >= 1 of the unit box), coarsest first, from the level-1 mesh `one`: each
// level is the previous one with the cells refined that the next finer cell
// set no longer holds, so nodes, transfer maps and patch grouping come from
// refine().
int adaptive_levels(const Mesh& one, const std::vector<uint64_t>& K, std::vector<Mesh>& out) {
  int top = 0;
  for (uint64_t c : K) top = std::max(top, ck_level(c));
  std::vector<std::vector<uint64_t>> sets(1, K);
  std::sort(sets[0].begin(), sets[0].end());
  for (int l = 1; l < top; ++l) {
    std::unordered_map<uint64_t, int> S;
    for (uint64_t c : sets.back()) S[c] = 1;
    std::vector<uint64_t> next;
    coarser_level(S, next);
    sets.push_back(next);
  }
  // sets[top-1] must be the level-1 mesh
  std::vector<uint64_t> K1;
  if (cell_keys(one, K1)) return -1;
  std::sort(K1.begin(), K1.end());
  if (K1 != sets.back()) return fail("adaptive levels: the coarsest level is not the level-1 mesh");
  out.assign(1, one);
  out[0].plev.clear();
  for (int l = top - 2; l >= 0; --l) {
    const Mesh& C = out.back();
    std::vector<uint64_t> KC;
    if (cell_keys(C, KC)) return -1;
    std::unordered_map<uint64_t, int> want;
    for (uint64_t c : sets[l]) want[c] = 1;
    std::vector<uint8_t> mark(C.ncells(), 0);
    for (int64_t c = 0; c < C.ncells(); ++c) mark[c] = want.count(KC[c]) ? 0 : 1;
    for (int64_t p = 0; p < C.ncells() / 8; ++p)
      for (int t = 1; t < 8; ++t)
        if (mark[p * 8 + t] != mark[p * 8])
          return fail("adaptive levels: a refinement splits a patch");
    Mesh F;
    if (refine(C, F, &mark)) return -1;
    std::vector<uint64_t> KF;
    if (cell_keys(F, KF)) return -1;
    std::sort(KF.begin(), KF.end());
    if (KF != sets[l]) return fail("adaptive levels: level %d is not reached by refinement", l);
    if (q2_hanging(F, F.hn_e, F.hn_f)) return -1;
    out.push_back(std::move(F));
  }
  return 0;
}

// Gascoigne's hierarchical mesh file (HierarchicalMesh3d::write_gup /
// read_gup, src/Mesh/hierarchicalmesh3d.cc:1534-1700; cell and boundary-cell
// text forms of src/Mesh/cell.h:182-192, boundarycell.h:96-108, hang lists of
// hanglistio.cc): the level-0 hexes and boundary quads become the macro
// mesh, the hexes without children the finest cells (as lattice keys of the
// unit box; coordinates are written with 6 significant digits and snapped).
int read_gup(const char* path, Mesh& base, std::vector<uint64_t>& leaves) {
  std::ifstream in(path);
  if (!in) return fail("cannot open %s", path);
  std::string sym;
  int dim = 0;
  int64_t nv = 0;
  in >> dim >> sym;
  if (!in || dim != 3 || sym != "dimension") return fail("%s: not a 3D .gup file", path);
  in >> nv >> sym;
  if (!in || nv <= 0 || sym != "vertexs") return fail("%s: bad vertex header", path);
  std::vector<double> X(nv * 3);
  for (int64_t i = 0; i < nv * 3; ++i) in >> X[i];
  int64_t nh = 0;
  in >> nh >> sym;
  if (!in || nh <= 0 || sym != "hexs") return fail("%s: bad hex header", path);
  std::vector<std::array<int32_t, 8>> hv(nh);
  std::vector<int32_t> hlev(nh), hnch(nh);
  for (int64_t h = 0; h < nh; ++h) {
    int64_t father, mat, nch;
    for (int m = 0; m < 8; ++m) in >> hv[h][m];
    in >> hlev[h] >> father >> mat >> sym >> nch;
    if (!in || sym != "@" || nch < 0 || nch > 8) return fail("%s: bad hex %lld", path, (long long)h);
    hnch[h] = (int32_t)nch;
    int64_t c;
    for (int64_t k = 0; k < nch; ++k) in >> c;
    in >> sym;
    if (sym != ":") return fail("%s: bad hex %lld", path, (long long)h);
    for (int k = 0; k < 6; ++k) in >> c;
    for (int m = 0; m < 8; ++m)
      if (hv[h][m] < 0 || hv[h][m] >= nv) return fail("%s: hex %lld: bad vertex", path, (long long)h);
  }
  // the face and edge hang lists: N vertices, "->", hanging, ":", 2 neighbours
  for (int n : {4, 2}) {
    int64_t cnt;
    in >> cnt >> sym;
    if (!in || sym != "hangs") return fail("%s: bad hang list", path);
    int64_t t;
    for (int64_t k = 0; k < cnt; ++k) {
      for (int a = 0; a < n; ++a) in >> t;
      in >> sym >> t >> sym >> t >> t;
    }
  }
  int64_t nq = 0;
  in >> nq >> sym;
  if (!in || sym != "boundaryquads") return fail("%s: bad boundary header", path);
  std::unordered_map<int32_t, int32_t> vmap;
  auto local = [&](int32_t v) {
    auto it = vmap.find(v);
    if (it != vmap.end()) return it->second;
    const int32_t id = (int32_t)vmap.size();
    vmap.emplace(v, id);
    for (int d = 0; d < 3; ++d) base.X.push_back(X[3 * (int64_t)v + d]);
    return id;
  };
  base = Mesh();
  base.dim = 3;
  base.npc = 8;
  for (int64_t h = 0; h < nh; ++h)
    if (hlev[h] == 0)
      for (int m = 0; m < 8; ++m) base.cells.push_back(local(hv[h][m]));
  for (int64_t k = 0; k < nq; ++k) {
    int64_t mat, lev, father, ofq, eiq, nch, c;
    int32_t q[4];
    in >> mat >> sym;
    for (int a = 0; a < 4; ++a) in >> q[a];
    in >> lev >> father >> ofq >> eiq >> sym >> nch;
    if (!in || sym != "@") return fail("%s: bad boundary quad %lld", path, (long long)k);
    for (int64_t a = 0; a < nch; ++a) in >> c;
    if (lev != 0) continue;
    for (int a = 0; a < 4; ++a) {
      if (q[a] < 0 || q[a] >= nv) return fail("%s: boundary quad %lld: bad vertex", path, (long long)k);
      base.bnd.push_back(local(q[a]));
    }
    base.bcol.push_back((int32_t)mat);
  }
  if (base.ncells() != 1) return fail("%s: the .gup reader takes the unit box (one macro hex)", path);
  for (int m = 0; m < 8; ++m)
    for (int d = 0; d < 3; ++d)
      if (base.X[3 * base.cells[m] + d] != (double)kHexCorner[m][d])
        return fail("%s: the macro hex is not the unit box in Gascoigne's corner order", path);
  // finest cells: snapped lattice keys
  leaves.clear();
  for (int64_t h = 0; h < nh; ++h) {
    if (hnch[h] != 0) continue;
    const double* o = &X[3 * (int64_t)hv[h][0]];
    const double* x1 = &X[3 * (int64_t)hv[h][6]];
    const double sz = x1[0] - o[0];
    if (!(sz > 0)) return fail("%s: hex %lld is not a box", path, (long long)h);
    const int L = (int)std::lround(-std::log2(sz));
    if (L < 1 || L > 18 || hlev[h] != L)
      return fail("%s: hex %lld: level %d does not match its size", path, (long long)h, hlev[h]);
    const double s = std::ldexp(1.0, -L);
    int64_t ijk[3];
    for (int d = 0; d < 3; ++d) ijk[d] = std::llround(o[d] / s);
    for (int m = 0; m < 8; ++m)
      for (int d = 0; d < 3; ++d) {
        const double want = (ijk[d] + kHexCorner[m][d]) * s;
        if (std::fabs(X[3 * (int64_t)hv[h][m] + d] - want) > 1e-4 * s)
          return fail("%s: hex %lld is not a cell of the dyadic unit box", path, (long long)h);
      }
    leaves.push_back(ckey(L, ijk[0], ijk[1], ijk[2]));
  }
  return 0;
}

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

int ghs_build_adaptive(const char* inp_path, int prerefine, int drop_coarse, int nadapt,
                       double radius, const double* centre, ghs_hier** out) {
  if (!out || !inp_path || !centre || prerefine < 1 || drop_coarse < 0 ||
      drop_coarse > prerefine || nadapt < 0 || nadapt > 2 || !(radius > 0))
    return fail("ghs_build_adaptive: bad arguments (nadapt 0..2, prerefine >= 1)");
  *out = nullptr;
  std::vector<Mesh> all(1);
  if (read_inp(inp_path, 3, all[0])) return -1;
  for (int r = 0; r < prerefine; ++r) {
    all.emplace_back();
    if (refine(all[r], all[r + 1])) return -1;
  }
  for (int s = 0; s < nadapt; ++s) {
    all.emplace_back();
    if (adapt(all[all.size() - 2], all.back(), radius, centre)) return -1;
  }
  if (nadapt > 0) {
    // the multigrid levels GascoigneMeshConstructor derives from the final
    // mesh (not the meshes after each step)
    std::vector<Mesh> lv;
    std::vector<uint64_t> K;
    if (cell_keys(all.back(), K) || adaptive_levels(all[1], K, lv)) return -1;
    all.resize(1);
    for (Mesh& m : lv) all.push_back(std::move(m));
  }
  ghs_hier* h = new ghs_hier();
  h->levels.assign(std::make_move_iterator(all.begin() + drop_coarse),
                   std::make_move_iterator(all.end()));
  *out = h;
  return 0;
}

int ghs_build_gup(const char* gup_path, int drop_coarse, ghs_hier** out) {
  if (!out || !gup_path || drop_coarse < 0) return fail("ghs_build_gup: bad arguments");
  *out = nullptr;
  std::vector<Mesh> all(2);
  std::vector<uint64_t> K;
  if (read_gup(gup_path, all[0], K)) return -1;
  if (refine(all[0], all[1])) return -1;
  std::vector<Mesh> lv;
  if (adaptive_levels(all[1], K, lv)) return -1;
  all.resize(1);
  for (Mesh& m : lv) all.push_back(std::move(m));
  if (drop_coarse >= (int)all.size()) return fail("ghs_build_gup: drop_coarse leaves no level");
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

int ghs_colour_order(int64_t n, const int32_t* rowptr, const int32_t* col, int32_t* perm,
                     int32_t* ncolours) {
  if (n < 0 || (n > 0 && (!rowptr || !col || !perm)) || !ncolours) {
    g_err = "ghs_colour_order: bad arguments";
    return -1;
  }
  // greedy colouring in node order: the smallest colour no coupled node
  // already holds
  std::vector<int32_t> colour(n, -1), seen;
  int32_t nc = 0;
  for (int64_t i = 0; i < n; ++i) {
    for (int32_t q = rowptr[i]; q < rowptr[i + 1]; ++q) {
      const int32_t j = col[q];
      if (j < 0 || j >= n) {
        g_err = "ghs_colour_order: column out of range";
        return -1;
      }
      const int32_t c = colour[j];
      if (c >= 0) {
        if ((int32_t)seen.size() <= c) seen.resize(c + 1, -1);
        seen[c] = (int32_t)i;
      }
    }
    int32_t c = 0;
    while (c < (int32_t)seen.size() && seen[c] == (int32_t)i) ++c;
    colour[i] = c;
    nc = std::max(nc, c + 1);
  }
  std::vector<int64_t> start(nc + 1, 0);
  for (int64_t i = 0; i < n; ++i) start[colour[i] + 1]++;
  for (int32_t c = 0; c < nc; ++c) start[c + 1] += start[c];
  for (int64_t i = 0; i < n; ++i) perm[start[colour[i]]++] = (int32_t)i;
  *ncolours = nc;
  return 0;
}

int ghs_level_patch_sizes(const ghs_hier* h, int l, int64_t* npatches, int64_t* nnz) {
  const Mesh* M = lev(h, l);
  if (!M) return -1;
  std::vector<int32_t> P;
  if (patches(*M, P)) return -1;
  const int nloc = M->dim == 3 ? 27 : 9;
  if (npatches) *npatches = (int64_t)P.size() / nloc;
  if (nnz) {
    std::vector<int32_t> rp, c;
    if (structure(M->n(), (int64_t)P.size() / nloc, nloc, P.data(), rp, c)) return -1;
    *nnz = rp.back();
  }
  return 0;
}

int ghs_level_patches(const ghs_hier* h, int l, int32_t* nodes) {
  const Mesh* M = lev(h, l);
  if (!M) return -1;
  std::vector<int32_t> P;
  if (patches(*M, P)) return -1;
  std::memcpy(nodes, P.data(), P.size() * sizeof(int32_t));
  return 0;
}

int ghs_level_patch_structure(const ghs_hier* h, int l, int32_t* rowptr, int32_t* col) {
  const Mesh* M = lev(h, l);
  if (!M) return -1;
  std::vector<int32_t> P, rp, c;
  if (patches(*M, P)) return -1;
  const int nloc = M->dim == 3 ? 27 : 9;
  if (structure(M->n(), (int64_t)P.size() / nloc, nloc, P.data(), rp, c)) return -1;
  std::memcpy(rowptr, rp.data(), rp.size() * sizeof(int32_t));
  if (col) std::memcpy(col, c.data(), c.size() * sizeof(int32_t));
  return 0;
}

int ghs_level_assemble_stokes(const ghs_hier* h, int l, double visc, double alpha, int ncolours,
                              const int32_t* colours, int dircomps, const double* u0, double* rhs,
                              double* val) {
  const Mesh* M = lev(h, l);
  if (!M) return -1;
  if (M->dim != 3) return fail("Stokes-LPS assembly is three-dimensional");
  if (!(visc > 0) || !(alpha >= 0)) return fail("Stokes-LPS: visc must be > 0 and alpha >= 0");
  if (dircomps & ~0xF) return fail("Stokes-LPS: Dirichlet component mask has bits beyond 4 components");
  if ((u0 == nullptr) != (rhs == nullptr)) return fail("Stokes-LPS: u0 and rhs go together");
  std::vector<int32_t> P, rp, c, d;
  if (patches(*M, P)) return -1;
  const int64_t n = M->n(), np = (int64_t)P.size() / 27;
  if (structure(n, np, 27, P.data(), rp, c)) return -1;
  // patch shapes (one per distinct box size; a uniform level has one)
  std::vector<int32_t> shape(np);
  std::vector<std::array<double, 3>> sizes;
  std::vector<std::unique_ptr<StokesPatch>> elem;
  for (int64_t p = 0; p < np; ++p) {
    double hb[3];
    if (patch_box(*M, &P[p * 27], hb)) return fail("patch %lld is not an axis-aligned box of 8 equal cells", (long long)p);
    int s = 0;
    while (s < (int)sizes.size() && !(sizes[s][0] == hb[0] && sizes[s][1] == hb[1] && sizes[s][2] == hb[2])) ++s;
    if (s == (int)sizes.size()) {
      if (s >= 64) return fail("Stokes-LPS: more than 64 distinct patch sizes");
      sizes.push_back({hb[0], hb[1], hb[2]});
      elem.emplace_back(new StokesPatch(hb[0], hb[1], hb[2], visc, alpha));
    }
    shape[p] = s;
  }
  // node -> (patch, local index), ascending patch order
  std::vector<int32_t> pstart(n + 1, 0), plist(np * 27);
  std::vector<int8_t> ploc(np * 27);
  for (int64_t k = 0; k < np * 27; ++k) pstart[P[k] + 1]++;
  for (int64_t i = 0; i < n; ++i) pstart[i + 1] += pstart[i];
  {
    std::vector<int32_t> fill(pstart.begin(), pstart.end() - 1);
    for (int64_t p = 0; p < np; ++p)
      for (int a = 0; a < 27; ++a) {
        const int32_t k = fill[P[p * 27 + a]]++;
        plist[k] = (int32_t)p;
        ploc[k] = (int8_t)a;
      }
  }
  dir_nodes(M, ncolours, colours, d);
  std::vector<uint8_t> isdir(n, 0);
  if (dircomps)
    for (int32_t i : d) isdir[i] = 1;
  int bad = 0;
#pragma omp parallel for schedule(dynamic, 1024) reduction(+ : bad)
  for (int64_t i = 0; i < n; ++i) {
    const int32_t rs = rp[i], re = rp[i + 1];
    double* vr = val + (int64_t)rs * 16;
    std::fill(vr, vr + (int64_t)(re - rs) * 16, 0.0);
    for (int32_t k = pstart[i]; k < pstart[i + 1]; ++k) {
      const int64_t p = plist[k];
      const int a = ploc[k];
      const StokesPatch& E = *elem[shape[p]];
      for (int b = 0; b < 27; ++b) {
        const int32_t j = P[p * 27 + b];
        const int32_t* it = std::lower_bound(c.data() + rs, c.data() + re, j);
        if (it == c.data() + re || *it != j) {
          bad++;
          continue;
        }
        double* e = val + (int64_t)(it - c.data()) * 16;
        for (int t = 0; t < 16; ++t) e[t] += E.blk[a][b][t];
      }
    }
    if (rhs) {
      // residual of the lifted state: rhs = -A u0 before the Dirichlet rows go
      double r[4] = {0, 0, 0, 0};
      for (int32_t q = rs; q < re; ++q) {
        const double* e = val + (int64_t)q * 16;
        const double* x = u0 + 4 * (int64_t)c[q];
        for (int s = 0; s < 4; ++s)
          for (int t = 0; t < 4; ++t) r[s] -= e[4 * s + t] * x[t];
      }
      for (int s = 0; s < 4; ++s)
        rhs[4 * i + s] = (isdir[i] && (dircomps >> s & 1)) ? 0.0 : r[s];
    }
    // SparseBlockMatrix::dirichlet (src/LinAlg/sparseblockmatrix.xx:205-218,
    // fmatrixblock.xx:54-82): rows and columns of the constrained components
    // zero, their diagonal entries one
    for (int32_t q = rs; q < re; ++q) {
      double* e = val + (int64_t)q * 16;
      for (int s = 0; s < 4; ++s) {
        if (!(dircomps >> s & 1)) continue;
        if (isdir[i])
          for (int t = 0; t < 4; ++t) e[4 * s + t] = 0.0;
        if (isdir[c[q]])
          for (int t = 0; t < 4; ++t) e[4 * t + s] = 0.0;
        if (isdir[i] && c[q] == i) e[4 * s + s] = 1.0;
      }
    }
  }
  if (bad) return fail("patch structure misses %d couplings", bad);
  return 0;
}

int ghs_level_assemble_q2(const ghs_hier* h, int l, int ncolours, const int32_t* colours,
                          double* val) {
  const Mesh* M = lev(h, l);
  if (!M) return -1;
  std::vector<int32_t> rp, c, d;
  std::vector<double> v;
  if (q2_matrix(M, rp, c, v)) return -1;
  const int64_t n = M->n();
  dir_nodes(M, ncolours, colours, d);
  std::vector<uint8_t> isdir(n, 0);
  for (int32_t i : d) isdir[i] = 1;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i)
    for (int32_t q = rp[i]; q < rp[i + 1]; ++q) {
      if (isdir[i])
        val[q] = (c[q] == i) ? 1.0 : 0.0;
      else if (isdir[c[q]])
        val[q] = 0.0;
      else
        val[q] = v[q];
    }
  return 0;
}

int ghs_level_hn_sizes(const ghs_hier* h, int l, int64_t* nedges, int64_t* nfaces) {
  const Mesh* M = lev(h, l);
  if (!M) return -1;
  if (nedges) *nedges = (int64_t)M->hn_e.size() / 4;
  if (nfaces) *nfaces = (int64_t)M->hn_f.size() / 10;
  return 0;
}

int ghs_level_hn(const ghs_hier* h, int l, int32_t* edges, int32_t* faces) {
  const Mesh* M = lev(h, l);
  if (!M) return -1;
  if (edges) std::memcpy(edges, M->hn_e.data(), M->hn_e.size() * sizeof(int32_t));
  if (faces) std::memcpy(faces, M->hn_f.data(), M->hn_f.size() * sizeof(int32_t));
  return 0;
}

int ghs_level_q2_condensed_nnz(ghs_hier* h, int l, int ncolours, const int32_t* colours,
                               int64_t* nnz) {
  const Mesh* M = lev(h, l);
  if (!M || !nnz) return fail("ghs_level_q2_condensed_nnz: bad arguments");
  std::vector<int32_t> rp, c, d;
  std::vector<double> v;
  if (q2_matrix(M, rp, c, v)) return -1;
  const int64_t n = M->n();
  dir_nodes(M, ncolours, colours, d);
  std::vector<uint8_t> isdir(n, 0);
  for (int32_t i : d) isdir[i] = 1;
  ghs_hier::Condensed& C = h->q2c[l];
  if (condense(n, rp, c, v, M->hn_e, M->hn_f, isdir, C.rp, C.col, C.val)) return -1;
  *nnz = C.rp[n];
  return 0;
}

int ghs_level_q2_condensed(ghs_hier* h, int l, int32_t* rowptr, int32_t* col, double* val) {
  auto it = h ? h->q2c.find(l) : h->q2c.end();
  if (!h || it == h->q2c.end()) return fail("ghs_level_q2_condensed: call ghs_level_q2_condensed_nnz first");
  const ghs_hier::Condensed& C = it->second;
  std::memcpy(rowptr, C.rp.data(), C.rp.size() * sizeof(int32_t));
  std::memcpy(col, C.col.data(), C.col.size() * sizeof(int32_t));
  std::memcpy(val, C.val.data(), C.val.size() * sizeof(double));
  h->q2c.erase(it);
  return 0;
}

int ghs_level_face_mass(const ghs_hier* h, int l, int colour, double* m) {
  const Mesh* M = lev(h, l);
  if (!M) return -1;
  if (M->dim != 3) return fail("face masses are three-dimensional");
  std::fill(m, m + M->n(), 0.0);
  for (size_t f = 0; f < M->bcol.size(); ++f) {
    if (M->bcol[f] != colour) continue;
    const int32_t* q = &M->bnd[4 * f];
    double e1 = 0, e2 = 0;
    for (int d = 0; d < 3; ++d) {
      const double a = M->X[3 * (int64_t)q[1] + d] - M->X[3 * (int64_t)q[0] + d];
      const double b = M->X[3 * (int64_t)q[3] + d] - M->X[3 * (int64_t)q[0] + d];
      e1 += a * a;
      e2 += b * b;
    }
    // bilinear hats on a rectangle: int phi_i ds = area / 4
    const double w = std::sqrt(e1) * std::sqrt(e2) / 4.0;
    for (int t = 0; t < 4; ++t) m[q[t]] += w;
  }
  return 0;
}

}  // extern "C"
