// Solve a uniformly refined Q1 Poisson problem through the C++ API surface
// (include/gascoigne_hip.hpp) and check it against the oracle: the same
// calls a Gascoigne CudaMultiLevelSolver makes -- LinearMg, Gmres, Solve --
// with the reference's parameters (V(2,2), damped Jacobi omega 0.8, GMRES(10)).
//
//   api_solve <inp> <dim> <prerefine> <ndirect> <prec f64|f32|f16> <krylov gmres|fgmres>
//             <order none|lexicographic>
//
// Checks (exit status 0 when all hold):
//   * FP64 with GMRES: LinearMg and Newton/GMRES iterates within 1e-12 (relative)
//     of the oracle's, iteration counts equal;
//   * any precision / Krylov method: the true FP64 residual reaches 1e-10
//     relative and the solution is within 1e-8 of the oracle's FP64 solve.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "gascoigne_hip.hpp"
#include "gascoigne_oracle.h"
#include "gh_structure.h"

namespace G = gascoigne_hip;

struct Level {
  int64_t n = 0;
  std::vector<int32_t> rowptr, col, dir, c2f, zw, vi, ac;
  std::vector<double> val, xyz;
};

static void sck(int rc) {
  if (rc) {
    std::fprintf(stderr, "gh_structure: %s\n", ghs_last_error());
    std::exit(2);
  }
}

static std::vector<Level> build(const char* inp, int dim, int pre) {
  ghs_hier* h = nullptr;
  sck(ghs_build(inp, dim, pre, 1, &h));
  const int nl = ghs_nlevels(h);
  std::vector<Level> L(nl);
  const int32_t colour = 1;
  for (int l = 0; l < nl; ++l) {
    Level& v = L[l];
    int64_t nc, nnz;
    sck(ghs_level_sizes(h, l, &v.n, &nc, &nnz));
    v.rowptr.resize(v.n + 1);
    v.col.resize(nnz);
    v.val.resize(nnz);
    v.xyz.resize(3 * v.n);
    sck(ghs_level_structure(h, l, v.rowptr.data(), v.col.data()));
    sck(ghs_level_assemble(h, l, nullptr, 1, &colour, v.val.data()));
    sck(ghs_level_coords(h, l, v.xyz.data()));
    int64_t nd;
    sck(ghs_level_dirichlet_count(h, l, 1, &colour, &nd));
    v.dir.resize(nd);
    sck(ghs_level_dirichlet(h, l, 1, &colour, v.dir.data()));
    if (l > 0) {
      int64_t ncs, n2, n4, n8;
      sck(ghs_level_transfer_sizes(h, l, &ncs, &n2, &n4, &n8));
      v.c2f.resize(ncs);
      v.zw.resize(3 * n2);
      v.vi.resize(5 * n4);
      v.ac.resize(9 * n8);
      sck(ghs_level_transfer(h, l, v.c2f.data(), v.zw.data(), v.vi.data(), v.ac.data()));
    }
  }
  ghs_destroy(h);
  return L;
}

// SplitMix64 uniform in [-1, 1)
static std::vector<double> seeded(uint64_t seed, int64_t n) {
  std::vector<double> f(n);
  for (int64_t k = 0; k < n; ++k) {
    uint64_t z = seed + (uint64_t)(k + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    f[k] = 2.0 * ((double)(z >> 11) * (1.0 / 9007199254740992.0)) - 1.0;
  }
  return f;
}

static double relerr(const std::vector<double>& a, const std::vector<double>& b) {
  double num = 0, den = 0;
  for (size_t i = 0; i < a.size(); ++i) {
    num += (a[i] - b[i]) * (a[i] - b[i]);
    den += b[i] * b[i];
  }
  return std::sqrt(num / den);
}

int main(int argc, char** argv) {
  if (argc != 8) {
    std::fprintf(stderr, "usage: api_solve inp dim prerefine ndirect prec krylov order\n");
    return 2;
  }
  const char* inp = argv[1];
  const int dim = std::atoi(argv[2]), pre = std::atoi(argv[3]), ndirect = std::atoi(argv[4]);
  const std::string prec = argv[5], krylov = argv[6], order = argv[7];
  std::vector<Level> L = build(inp, dim, pre);
  const int nl = (int)L.size();
  Level& F = L.back();
  std::vector<double> f = seeded(7, F.n);
  for (int32_t d : F.dir) f[d] = 0.0;

  // ---------------------------------------------------------------- device
  G::Context ctx(0);
  std::vector<G::CudaCSRMatrixInterface> A;
  std::vector<G::CudaMgInterpolatorNested> T;
  std::vector<G::DirichletMask> M;
  A.reserve(nl);
  T.reserve(nl);
  M.reserve(nl);
  std::vector<const G::CudaCSRMatrixInterface*> pa;
  std::vector<const G::CudaMgInterpolatorNested*> pt;
  std::vector<const G::DirichletMask*> pm;
  for (int l = 0; l < nl; ++l) {
    // an ordered solver builds its own device matrices: hand it host-only ones
    if (order == "lexicographic")
      A.emplace_back(G::CudaCSRMatrixInterface::HostOnly{}, ctx, L[l].n, L[l].rowptr, L[l].col,
                     L[l].val);
    else
      A.emplace_back(ctx, L[l].n, L[l].rowptr, L[l].col, L[l].val);
    pa.push_back(&A.back());
    if (l > 0) {
      T.emplace_back(ctx, L[l].c2f, L[l].zw, L[l].vi, L[l].ac);
      pt.push_back(&T.back());
    } else {
      pt.push_back(nullptr);
    }
    M.emplace_back(ctx, L[l].n, 1, L[l].dir);
    pm.push_back(&M.back());
  }
  G::MgParams P;
  P.p.mgtype = 'V';
  P.p.iterpre = 2;
  P.p.iterpost = 2;
  P.p.iterexact = 10;
  P.p.omega = 0.8;
  P.p.mgomega = 1.0;
  P.p.ndirect = ndirect;
  P.p.prec = prec == "f64" ? GH_F64 : prec == "f32" ? GH_F32 : GH_F16;
  P.p.krylov = krylov == "fgmres" ? 1 : 0;
  std::vector<std::vector<int32_t>> ord;
  if (order == "lexicographic") {
    for (int l = 0; l < nl; ++l) {
      std::vector<int32_t> o(L[l].n);
      std::iota(o.begin(), o.end(), 0);
      const double* x = L[l].xyz.data();
      std::stable_sort(o.begin(), o.end(), [&](int32_t a, int32_t b) {
        for (int d = 2; d >= 0; --d)
          if (x[3 * a + d] != x[3 * b + d]) return x[3 * a + d] < x[3 * b + d];
        return false;
      });
      ord.push_back(std::move(o));
    }
  }
  G::CudaMultiLevelSolver mg(ctx, pa, pt, {}, pm, P, ord.empty() ? nullptr : &ord);

  // ---------------------------------------------------------------- oracle
  or_mg* om = or_mg_create(nl);
  om->mgtype = 'V';
  om->iterpre = 2;
  om->iterpost = 2;
  om->iterexact = 10;
  om->omega = 0.8;
  om->mgomega = 1.0;
  om->smoother = 0;
  const int32_t dc = 0;
  std::vector<or_transfer*> ot(nl, nullptr);
  for (int l = 0; l < nl; ++l) {
    or_level& Lv = om->L[l];
    if (or_matrix_init(&Lv.A, (int)L[l].n, 1, L[l].rowptr.data(), L[l].col.data(),
                       L[l].val.data()))
      return 2;
    if (l > 0) {
      ot[l] = or_transfer_create((int)L[l].c2f.size(), L[l].c2f.data(), (int)L[l].zw.size() / 3,
                                 L[l].zw.data(), (int)L[l].vi.size() / 5, L[l].vi.data(),
                                 (int)L[l].ac.size() / 9, L[l].ac.data());
      Lv.T = ot[l];
    }
    Lv.hn.kind = dim == 3 ? 13 : 12;
    Lv.ndir = (int)L[l].dir.size();
    Lv.dir = L[l].dir.data();
    Lv.ndircomp = 1;
    Lv.dircomp = &dc;
    Lv.direct = L[l].n < ndirect ? 1 : 0;
  }
  if (or_mg_setup(om)) return 2;

  bool ok = true;
  const bool faithful = prec == "f64" && krylov == "gmres";
  // LinearMg from zero (FP64 only: the oracle runs the same iteration)
  if (prec == "f64") {
    G::CudaVectorInterface U(ctx, F.n), Fv(ctx, f);
    G::CGInfo info(1e-8, 1e-30, 30);
    mg.LinearMg(nl - 1, 0, U, Fv, info);
    std::vector<double> uo(F.n, 0.0);
    or_cginfo oi;
    or_cginfo_init(&oi, 1e-8, 1e-30, 30);
    or_linear_mg(om, nl - 1, 0, uo.data(), f.data(), &oi);
    const double e = relerr(U.copy_back(), uo);
    std::printf("LinearMg: iterations %d (oracle %d), status %s, rel. diff %.3e\n",
                info.iteration(), oi.iteration, info.status().c_str(), e);
    ok = ok && info.iteration() == oi.iteration && e <= 1e-12;
  }
  // Solve: Newton / iterative refinement with GMRES(10) or flexible GMRES(10)
  G::CudaVectorInterface U(ctx, F.n), Fv(ctx, f);
  G::CGInfo linfo(krylov == "fgmres" ? 0.5e-10 : 1e-10, 1e-30, 100);
  G::NLInfo nlinfo(1e-10, 1e-30, 50);
  const gh_solve_report rep = mg.Solve(U, Fv, linfo, nlinfo, 10, 1);
  const std::vector<double> u = U.copy_back();
  std::vector<double> uo(F.n, 0.0);
  or_cginfo ol;
  or_nlinfo on;
  or_cginfo_init(&ol, 1e-10, 1e-30, 100);
  or_nlinfo_init(&on, 1e-10, 1e-30, 50);
  or_solve_report orep;
  or_newton_solve(om, 0, 10, uo.data(), f.data(), &ol, &on, 1, &orep);
  std::vector<double> r(F.n);
  or_residual(&om->L[nl - 1].A, r.data(), u.data(), f.data());
  for (int32_t d : F.dir) r[d] = 0.0;
  const double rr = or_norm(r.data(), F.n) / rep.first_residual;
  const double e = relerr(u, uo);
  std::printf("Solve: status %s, newton %d (oracle %d), gmres %d (oracle %d), "
              "true rel. residual %.3e, rel. diff vs oracle %.3e\n",
              nlinfo.status().c_str(), rep.newton_iterations, orep.newton_iterations,
              rep.gmres_iterations, orep.gmres_iterations, rr, e);
  ok = ok && nlinfo.status() == "converged" && rr <= 1e-10 * 1.0001;
  if (faithful)
    ok = ok && e <= 1e-12 && rep.newton_iterations == orep.newton_iterations &&
         std::abs(rep.gmres_iterations - orep.gmres_iterations) <= 1;
  else
    ok = ok && e <= 1e-8;
  for (int l = 0; l < nl; ++l) or_matrix_free(&om->L[l].A);
  for (or_transfer* t : ot)
    if (t) or_transfer_free(t);
  or_mg_free(om);
  std::printf("%s\n", ok ? "PASS" : "FAIL");
  return ok ? 0 : 1;
}
