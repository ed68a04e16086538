"""The structure producer against the compiled reference at the workloads'
full sizes, through digests (tests/golden/digests_*.json, written by
tests/golden/make_digests.py from Gascoigne's own MeshAgent, CGDisc and
NavierStokesLps; the arrays themselves are hundreds of MB).

Every level of the 128^3 Q1 Poisson hierarchy (config 2), of the 64^3
Stokes-LPS channel (config 5) and of the Q2 hierarchy with two adaptive
refinements in a ball (config 4: MeshAgent::refine_nodes, the multigrid
levels GascoigneMeshConstructor derives, Q2 hanging nodes): renumbered by
node coordinates, the CSR pattern, the cells, the Dirichlet nodes, the
mesh-transfer maps (c2f, 2er, 4er, 8er) and the hanging-node constraints
have the reference's SHA-256; the matrix values agree through weighted sums
to 1e-12 (the assemblies differ by rounding only)."""
import json
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import digests as D  # noqa: E402

from gascoignelib_amd.structure import (q2_hierarchy, stokes_hierarchy,  # noqa: E402
                                        uniform_hierarchy)


def load(name):
    with open(os.path.join(HERE, "golden", f"digests_{name}.json")) as f:
        return json.load(f)


def compare(ref, hier, mesh):
    nl = ref["nlevels"]
    assert hier.nlevels == nl
    ranks = []
    for s in range(nl):
        m = nl - 1 - s  # the reference numbers levels from the finest
        lev = hier.level(s)
        rc = ref["levels"][m]["canonical"]
        mine, rank = D.canonical_level(lev.coords, lev.rowptr, lev.col, lev.val,
                                       mesh.cells(s), lev.dirichlet)
        ranks.append(rank)
        assert mine["nnodes"] == rc["nnodes"]
        assert mine["ncells"] == rc["ncells"] and mine["cells_sha256"] == rc["cells_sha256"]
        assert mine["ndirichlet"] == rc["ndirichlet"]
        assert mine["dirichlet_sha256"] == rc["dirichlet_sha256"]
        a, b = mine["csr"], rc["csr"]
        assert (a["n"], a["nnz"]) == (b["n"], b["nnz"])
        assert a["pattern_sha256"] == b["pattern_sha256"]
        scale = b["values"]["sum_abs"]
        for k in ("sum", "sum_abs", "weighted"):
            assert abs(a["values"][k] - b["values"][k]) <= 1e-12 * 17 * 16 * scale, k
        assert abs(a["values"]["sum_sq"] - b["values"]["sum_sq"]) <= 1e-12 * b["values"]["sum_sq"]
        if "hn_edge_sha256" in rc:
            hn = D.canonical_hn(lev.hn_edges, lev.hn_faces, rank)
            assert (hn["nhn_edge"], hn["nhn_face"]) == (rc["nhn_edge"], rc["nhn_face"])
            assert hn["hn_edge_sha256"] == rc["hn_edge_sha256"]
            assert hn["hn_face_sha256"] == rc["hn_face_sha256"]
    for s in range(1, nl):
        m = nl - 1 - s
        t = D.canonical_transfer(*hier.level(s).transfer, ranks[s], ranks[s - 1])
        assert t == ref["levels"][m]["canonical"]["transfer_from_coarser"]


def test_q1_poisson_128_equals_reference():
    ref = load("q1_3d_128")
    h = uniform_hierarchy("cube.inp", 3, 7, drop_coarse=1)
    compare(ref, h, h.mesh)


def test_stokes_channel_64_equals_reference():
    ref = load("stokes_3d_64")
    h = stokes_hierarchy(6, drop_coarse=1)
    compare(ref, h, h.mesh)


def test_q2_adaptive_64_equals_reference():
    ref = load("q2_3d_64_ada")
    r, cx, cy, cz, steps = ref["sphere"]
    h = q2_hierarchy(6, adaptive=int(steps), radius=r, centre=(cx, cy, cz))
    compare(ref, h, h.mesh)
