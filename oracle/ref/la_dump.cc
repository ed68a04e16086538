// Fixture dumper for the linear-algebra building blocks of the reference CPU
// path, run on Gascoigne's own mesh hierarchy.  Everything below the driver
// glue is reference code compiled from the reference sources in place
// (oracle/ref/Makefile):
//   * CGDisc<DIM,DEG>::Structure / Matrix for the Laplace equation
//     (src/DofHandler/cgdisc.h:299-315, 599-635; src/Problems/laplace.h)
//   * SparseBlockMatrix::dirichlet / vmult / Jacobi
//     (src/LinAlg/sparseblockmatrix.xx:47-56, 205-218, 378-392)
//   * MgInterpolatorMatrix::BasicInit / restrict_zero / prolongate_add
//     (src/Discretization/Q1/mginterpolatormatrix.cc:34-147)
//   * HNStructureQ1x / Q2x Average / Distribute / Zero via CGDisc::HN*
//   * nvector::norm / operator* (src/Common/nvector.h)
//   * SparseBlockIlu<FMatrixBlock<1>>: ConstructStructure / copy_entries /
//     compute_ilu / solve (src/LinAlg/sparseblockilu.xx:70-246) on levels of
//     at most 1100 nodes, with the identity order and the reversed order
//   * disc Q13dLps: NavierStokesLps<3> (src/Problems/navierstokes.h:92-270,
//     parameters from the param file's Equation block) linearised at U = 0 on
//     CGDiscQ13dLps (cgdisc.h:1314-1319), 4x4 blocks (p, u, v, w), on every
//     level with patches: structure and block values only
// The strong-Dirichlet loop mirrors CGDisc::StrongDirichletMatrix
// (cgdisc.h:1139-1150) for colour list `colours` and component 0 (the
// velocity components 1..3 for Q13dLps), since the ProblemDescriptor plumbing
// lives in the unbuildable solver sources.
//
// usage: la_dump <paramfile> <outdir> <disc Q12d|Q13d|Q22d|Q23d|Q13dLps> <colours,...>
//        [r cx cy cz nsteps]   (adaptive node refinement inside a sphere)
#include "cgdisc.h"
#include "gascoignemesh.h"
#include "laplace.h"
#include "navierstokes.h"
#include "meshagent.h"
#include "mginterpolatormatrix.h"
#include "npy.h"
#include "sparseblockmatrix.h"
#include "fmatrixblock.h"
#include "sparsestructure.h"
#include "sparseblockilu.h"
#include "columndiagstencil.h"

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <sstream>
#include <string>
#include <vector>

using namespace Gascoigne;

// distance of node i from (cx, cy, cz); 2D meshes ignore cz (the
// AdaptiveElasticity2d examples refine the nodes inside a circle)
static double node_distance(const GascoigneMesh* M, int i, double cx, double cy, double cz) {
  if (M->dimension() == 2) {
    const Vertex2d& x = M->vertex2d(i);
    return std::sqrt((x.x() - cx) * (x.x() - cx) + (x.y() - cy) * (x.y() - cy));
  }
  const Vertex3d& x = M->vertex3d(i);
  return std::sqrt((x.x() - cx) * (x.x() - cx) + (x.y() - cy) * (x.y() - cy) +
                   (x.z() - cz) * (x.z() - cz));
}

// SplitMix64 -> uniform double in [-1,1): v = 2*((x>>11)*2^-53) - 1
struct SplitMix64 {
  uint64_t s;
  explicit SplitMix64(uint64_t seed) : s(seed) {}
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform_pm1() { return 2.0 * ((double)(next() >> 11) * (1.0 / 9007199254740992.0)) - 1.0; }
};

static GlobalVector seeded(int n, uint64_t seed) {
  GlobalVector v(1, n);
  SplitMix64 r(seed);
  for (int i = 0; i < n; ++i) v(i, 0) = r.uniform_pm1();
  return v;
}

static void save_gv(const std::string& p, const GlobalVector& v) {
  std::vector<double> a(v.begin(), v.end());
  npy::save(p, a);
}

template <class DISC, int DIM>
static void dump(MeshAgent& MA, const ParamFile& pf, const std::vector<int>& colours,
                 const std::string& out) {
  const int nl = MA.nlevels();
  std::vector<GlobalVector> xs(nl);
  for (int l = 0; l < nl; ++l) {
    const GascoigneMesh* M = MA.GetMesh(l);
    DISC disc;
    disc.BasicInit(pf);
    disc.ReInit(M);
    const int n = disc.ndofs();
    const std::string p = out + "/L" + std::to_string(l) + "_";

    SparseStructure S;
    disc.Structure(&S);
    SparseBlockMatrix<FMatrixBlock<1>> A;
    A.ReInit(&S);
    A.zero();
    GlobalVector u(1, n);
    u.zero();
    LaplaceData ld;
    ld.visc = 1.0;
    Laplace<DIM> eq(ld);
    disc.Matrix(A, u, eq, 1.0);
    std::vector<int> comp0(1, 0);
    for (int col : colours) {
      const IntVector& bv = *M->VertexOnBoundary(col);
      for (size_t i = 0; i < bv.size(); ++i) A.dirichlet(bv[i], comp0);
    }

    const ColumnDiagStencil* ST = dynamic_cast<const ColumnDiagStencil*>(A.GetStencil());
    std::vector<int32_t> rowptr(ST->start().begin(), ST->start().end());
    std::vector<int32_t> col(ST->col().begin(), ST->col().end());
    std::vector<int32_t> diag(ST->diag().begin(), ST->diag().end());
    std::vector<double> val(A.size());
    for (int pos = 0; pos < A.size(); ++pos) val[pos] = (*A.mat(pos))(0, 0);
    npy::save(p + "rowptr.npy", rowptr);
    npy::save(p + "col.npy", col);
    npy::save(p + "diag.npy", diag);
    npy::save(p + "val.npy", val);

    // golden vectors for SpMV / residual / Jacobi / norms
    GlobalVector x = seeded(n, 1000 + l), b = seeded(n, 2000 + l);
    save_gv(p + "x.npy", x);
    save_gv(p + "b.npy", b);
    GlobalVector y(1, n);
    y.zero();
    A.vmult(y, x, 1.0);
    save_gv(p + "Ax.npy", y);
    GlobalVector h(1, n);
    h.equ(1.0, b);
    A.vmult(h, x, -1.0);
    save_gv(p + "res.npy", h);
    A.Jacobi(h);
    save_gv(p + "jac.npy", h);
    double nd[2] = {x.norm(), x * b};
    npy::save(p + "norm_dot.npy", nd, {2});

    // ILU(0) in a given node order: factors and one solve of a seeded vector
    if (n <= 1100) {
      GlobalVector s0 = seeded(n, 5000 + l);
      save_gv(p + "ilu_in.npy", s0);
      for (int kind = 0; kind < 2; ++kind) {
        nvector<int> perm(n);
        for (int i = 0; i < n; ++i) perm[i] = kind == 0 ? i : n - 1 - i;
        SparseBlockIlu<FMatrixBlock<1>> I;
        I.ConstructStructure(perm, A);
        I.zero();
        I.copy_entries(A);
        I.compute_ilu();
        const ColumnDiagStencil* IS = dynamic_cast<const ColumnDiagStencil*>(I.GetStencil());
        std::vector<int32_t> irp(IS->start().begin(), IS->start().end());
        std::vector<int32_t> icol(IS->col().begin(), IS->col().end());
        std::vector<int32_t> idiag(IS->diag().begin(), IS->diag().end());
        std::vector<double> ival(I.size());
        for (int pos = 0; pos < I.size(); ++pos) ival[pos] = (*I.mat(pos))(0, 0);
        const std::string q = p + "ilu" + std::to_string(kind) + "_";
        npy::save(q + "rowptr.npy", irp);
        npy::save(q + "col.npy", icol);
        npy::save(q + "diag.npy", idiag);
        npy::save(q + "val.npy", ival);
        GlobalVector s1(s0);
        I.solve(s1);
        save_gv(q + "solve.npy", s1);
      }
    }

    // hanging-node operators on a seeded vector
    if (disc.nhanging() > 0) {
      GlobalVector a = seeded(n, 3000 + l);
      GlobalVector c(a), d(a);
      disc.HNAverage(a);
      disc.HNDistribute(c);
      disc.HNZero(d);
      save_gv(p + "hn_in.npy", seeded(n, 3000 + l));
      save_gv(p + "hn_average.npy", a);
      save_gv(p + "hn_distribute.npy", c);
      save_gv(p + "hn_zero.npy", d);
    }
    xs[l] = seeded(n, 4000 + l);
    save_gv(p + "t_in.npy", xs[l]);
  }
  // transfer l (fine, mesh level l) <-> l+1 (coarse)
  for (int l = 0; l + 1 < nl; ++l) {
    MgInterpolatorMatrix I;
    I.BasicInit(MA.GetTransfer(l));
    const int nf = xs[l].n(), nc = xs[l + 1].n();
    GlobalVector uc(1, nc), uf(1, nf);
    uc.zero();
    uf.zero();
    I.restrict_zero(uc, xs[l]);
    I.prolongate_add(uf, xs[l + 1]);
    const std::string p = out + "/L" + std::to_string(l) + "_";
    save_gv(p + "restrict.npy", uc);
    save_gv(p + "prolongate.npy", uf);
  }
}

// NavierStokesLps<3> at U = 0 on the patch levels: rowptr, col and the 4x4
// blocks (row-major) after the velocity Dirichlet rows of `colours`
static void dump_stokes(MeshAgent& MA, const ParamFile& pf, const std::vector<int>& colours,
                        const std::string& out) {
  NavierStokesData nd;
  nd.BasicInit(pf);
  NavierStokesLps<3> eq(nd);
  for (int l = 0; l < MA.nlevels(); ++l) {
    const GascoigneMesh* M = MA.GetMesh(l);
    if (!M->HasPatch()) continue;
    CGDiscQ13dLps disc;
    disc.BasicInit(pf);
    disc.ReInit(M);
    const int n = disc.ndofs();
    const std::string p = out + "/L" + std::to_string(l) + "_";
    SparseStructure S;
    disc.Structure(&S);
    SparseBlockMatrix<FMatrixBlock<4>> A;
    A.ReInit(&S);
    A.zero();
    GlobalVector u(4, n);
    u.zero();
    disc.Matrix(A, u, eq, 1.0);
    std::vector<int> vel = {1, 2, 3};
    for (int col : colours) {
      const IntVector& bv = *M->VertexOnBoundary(col);
      for (size_t i = 0; i < bv.size(); ++i) A.dirichlet(bv[i], vel);
    }
    const ColumnDiagStencil* ST = dynamic_cast<const ColumnDiagStencil*>(A.GetStencil());
    std::vector<int32_t> rowptr(ST->start().begin(), ST->start().end());
    std::vector<int32_t> col(ST->col().begin(), ST->col().end());
    std::vector<double> val((size_t)A.size() * 16);
    for (int pos = 0; pos < A.size(); ++pos)
      for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) val[(size_t)pos * 16 + r * 4 + c] = (*A.mat(pos))(r, c);
    npy::save(p + "stokes_rowptr.npy", rowptr);
    npy::save(p + "stokes_col.npy", col);
    npy::save(p + "stokes_val.npy", val.data(), {(size_t)A.size(), (size_t)16});
  }
}

int main(int argc, char** argv) {
  if (argc < 5) {
    std::fprintf(stderr, "usage: %s paramfile outdir disc colours [r cx cy cz nsteps]\n", argv[0]);
    return 2;
  }
  ParamFile pf(argv[1]);
  std::string out = argv[2], disc = argv[3];
  std::vector<int> colours;
  {
    std::stringstream ss(argv[4]);
    std::string t;
    while (std::getline(ss, t, ',')) colours.push_back(std::atoi(t.c_str()));
  }
  MeshAgent MA;
  MA.BasicInit(pf);
  if (argc >= 10) {
    double r = std::atof(argv[5]);
    double cx = std::atof(argv[6]), cy = std::atof(argv[7]), cz = std::atof(argv[8]);
    int steps = std::atoi(argv[9]);
    for (int s = 0; s < steps; ++s) {
      const GascoigneMesh* M = MA.GetMesh(0);
      IntVector ref;
      for (int i = 0; i < M->nnodes(); ++i)
        if (node_distance(M, i, cx, cy, cz) < r) ref.push_back(i);
      MA.refine_nodes(ref);
    }
  }
  if (disc == "Q12d") dump<CGDiscQ12d, 2>(MA, pf, colours, out);
  else if (disc == "Q22d") dump<CGDiscQ22d, 2>(MA, pf, colours, out);
  else if (disc == "Q13d") dump<CGDiscQ13d, 3>(MA, pf, colours, out);
  else if (disc == "Q23d") dump<CGDiscQ23d, 3>(MA, pf, colours, out);
  else if (disc == "Q13dLps") dump_stokes(MA, pf, colours, out);
  else {
    std::fprintf(stderr, "unknown disc %s\n", disc.c_str());
    return 2;
  }
  return 0;
}
