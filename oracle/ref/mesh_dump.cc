// Fixture dumper: builds Gascoigne's own multigrid mesh hierarchy with the
// reference MeshAgent (compiled from the reference sources by
// oracle/ref/Makefile) and writes, per level, the node coordinates, the cell
// and patch connectivity, the Dirichlet colour lists, the hanging-node maps
// and the mesh-transfer maps (c2f / 2er / 4er / 8er) as .npy files, and the
// hierarchical mesh itself as mesh.gup (MeshAgent::write_gup).
//
// Test infrastructure: the output is the bit-exact numbering that the
// structure producer in gascoignelib_amd/ must reproduce.
//
// Follows: MeshAgent::BasicInit / ReInit (src/Mesh/meshagent.cc:64-147,
// 361-406), GascoigneMeshTransfer (src/Mesh/gascoignemeshtransfer.h:35-73),
// HangingIndexHandler (src/Mesh/hangingindexhandler.h:94-103),
// BoundaryIndexHandler (src/Mesh/boundaryindexhandler.h).
//
// usage: mesh_dump <paramfile> <outdir> [refine_sphere r cx cy cz nsteps]
#include "meshagent.h"
#include "gascoignemesh.h"
#include "gascoignemeshtransfer.h"
#include "npy.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
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

static void dump_level(const MeshAgent& MA, int l, const std::string& out) {
  const GascoigneMesh* M = MA.GetMesh(l);
  const int dim = M->dimension();
  const int n = M->nnodes();
  const int nc = M->ncells();
  const std::string p = out + "/L" + std::to_string(l) + "_";

  std::vector<double> xyz((size_t)n * dim);
  for (int i = 0; i < n; ++i) {
    if (dim == 2) {
      xyz[2 * i] = M->vertex2d(i).x();
      xyz[2 * i + 1] = M->vertex2d(i).y();
    } else {
      xyz[3 * i] = M->vertex3d(i).x();
      xyz[3 * i + 1] = M->vertex3d(i).y();
      xyz[3 * i + 2] = M->vertex3d(i).z();
    }
  }
  npy::save(p + "coords.npy", xyz.data(), {(size_t)n, (size_t)dim});

  const int npc = M->nodes_per_cell(0);
  std::vector<int32_t> cells((size_t)nc * npc);
  for (int c = 0; c < nc; ++c)
    for (int k = 0; k < npc; ++k) cells[(size_t)c * npc + k] = M->vertex_of_cell(c, k);
  npy::save(p + "cells.npy", cells.data(), {(size_t)nc, (size_t)npc});

  if (M->HasPatch()) {
    const int np = M->npatches();
    const int npp = M->nodes_per_patch();
    std::vector<int32_t> pat((size_t)np * npp);
    for (int q = 0; q < np; ++q) {
      const IntVector* iv = M->IndicesOfPatch(q);
      for (int k = 0; k < npp; ++k) pat[(size_t)q * npp + k] = (*iv)[k];
    }
    npy::save(p + "patches.npy", pat.data(), {(size_t)np, (size_t)npp});
  }

  const BoundaryIndexHandler& BH = M->GetBoundaryIndexHandler();
  for (int col : BH.GetColors()) {
    const IntVector& v = BH.Verteces(col);
    std::vector<int32_t> vv(v.begin(), v.end());
    npy::save(p + "bnd" + std::to_string(col) + ".npy", vv);
  }

  const HangingIndexHandler& HH = M->GetHangingIndexHandler();
  {
    const auto* e = HH.GetStructure();
    std::vector<int32_t> a;
    for (const auto& kv : *e) {
      a.push_back(kv.first);
      for (int k = 0; k < 3; ++k) a.push_back(kv.second[k]);
    }
    npy::save(p + "hn_edge.npy", a.data(), {e->size(), (size_t)4});
  }
  {
    const auto* f = HH.GetStructureFace();
    std::vector<int32_t> a;
    for (const auto& kv : *f) {
      a.push_back(kv.first);
      for (int k = 0; k < 9; ++k) a.push_back(kv.second[k]);
    }
    npy::save(p + "hn_face.npy", a.data(), {f->size(), (size_t)10});
  }

  if (l + 1 < MA.nlevels()) {
    const GascoigneMeshTransfer* T = MA.GetTransfer(l);
    const IntVector& c2f = T->GetC2f();
    std::vector<int32_t> a(c2f.begin(), c2f.end());
    npy::save(p + "c2f.npy", a);
    std::vector<int32_t> z, v, acht;
    for (const auto& kv : T->GetZweier()) {
      z.push_back(kv.first);
      for (int k : kv.second) z.push_back(k);
    }
    for (const auto& kv : T->GetVierer()) {
      v.push_back(kv.first);
      for (int k : kv.second) v.push_back(k);
    }
    for (const auto& kv : T->GetAchter()) {
      acht.push_back(kv.first);
      for (int k : kv.second) acht.push_back(k);
    }
    npy::save(p + "zweier.npy", z.data(), {z.size() / 3, (size_t)3});
    npy::save(p + "vierer.npy", v.data(), {v.size() / 5, (size_t)5});
    npy::save(p + "achter.npy", acht.data(), {acht.size() / 9, (size_t)9});
  }
}

int main(int argc, char** argv) {
  if (argc < 3) {
    std::fprintf(stderr, "usage: %s paramfile outdir [r cx cy cz nsteps]\n", argv[0]);
    return 2;
  }
  ParamFile pf(argv[1]);
  std::string out = argv[2];
  MeshAgent MA;
  MA.BasicInit(pf);
  if (argc >= 8) {
    // adaptive refinement of the nodes inside a sphere, as the
    // AdaptiveElasticity3d examples drive MeshAgent::refine_nodes
    double r = std::atof(argv[3]);
    double cx = std::atof(argv[4]), cy = std::atof(argv[5]), cz = std::atof(argv[6]);
    int steps = std::atoi(argv[7]);
    for (int s = 0; s < steps; ++s) {
      const GascoigneMesh* M = MA.GetMesh(0);
      IntVector ref;
      for (int i = 0; i < M->nnodes(); ++i)
        if (node_distance(M, i, cx, cy, cz) < r) ref.push_back(i);
      MA.refine_nodes(ref);
    }
  }
  // the hierarchical mesh in Gascoigne's own text format (mesh.gup)
  MA.write_gup(out + "/mesh");
  std::FILE* f = std::fopen((out + "/meta.txt").c_str(), "w");
  std::fprintf(f, "nlevels %d\ndim %d\n", MA.nlevels(), MA.GetDimension());
  for (int l = 0; l < MA.nlevels(); ++l)
    std::fprintf(f, "level %d nnodes %d ncells %d\n", l, MA.GetMesh(l)->nnodes(),
                 MA.GetMesh(l)->ncells());
  std::fclose(f);
  for (int l = 0; l < MA.nlevels(); ++l) dump_level(MA, l, out);
  return 0;
}
