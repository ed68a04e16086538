"""Q2 patch structure and assembly of the structure producer against the
reference's own output (CPU).

The golden fixture q2_3d_8_ada was written by the compiled reference
(CGDiscQ23d with Laplace<3> on cube.inp, prerefine 3, patchdepth 1, Dirichlet
colour 1): its two coarsest levels are uniform (4^3 and 8^3 cells), so the
producer's matrices and patches of those meshes must equal the reference's
once nodes are matched by their coordinates (the producer numbers new nodes
in its own order).  Checked: CSR pattern, every value (to rounding: the
producer forms the exact integrals as tensor products, the reference sums a
3x3x3 Gauss rule), Dirichlet rows, and the patch node lists in the
reference's patch-local order.

The adaptive levels (MeshAgent::refine_nodes of the nodes inside a ball,
the multigrid levels GascoigneMeshConstructor derives from the refined mesh,
Q2 hanging nodes, the operator condensed over them) equal the reference's on
the fixture's refinement (two steps in a sphere on 8^3 cells): under the
coordinate match every level's nodes, cells and transfer maps, the hanging
edge constraints row for row and the hanging face constraints as (parent,
weight) sets (some faces list the 3x3 parents transposed; the weights are
symmetric).  Pinned further by what the discretisation fixes: the
constraints interpolate quadratics exactly, the condensed operator equals an
independent scipy T^T A T, is symmetric, annihilates constants and is exact
on quadratics; a ball holding the whole domain reproduces the uniform
refinement entry for entry; the oracle's HNStructureQ23d solve converges on
them.
"""
import os

import numpy as np
import pytest
import scipy.sparse as sp

from gascoignelib_amd.structure import q2_hierarchy

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def coord_keys(X):
    return [tuple(np.round(x, 9)) for x in X]


@pytest.fixture(scope="module")
def ref():
    return np.load(os.path.join(GOLDEN, "q2_3d_8_ada.npz"))


@pytest.mark.parametrize("level,mesh_level", [(1, 3), (2, 2)])
def test_q2_matrix_matches_reference(ref, level, mesh_level):
    h = q2_hierarchy(3)
    lev = h.levels[level]
    p = f"L{mesh_level}_"
    X = ref[p + "coords"]
    assert len(X) == lev.n
    pos = {k: i for i, k in enumerate(coord_keys(lev.coords))}
    perm = np.array([pos[k] for k in coord_keys(X)])  # reference node -> producer node
    assert len(set(perm)) == lev.n
    R = sp.csr_matrix((ref[p + "val"], ref[p + "col"], ref[p + "rowptr"]), shape=(lev.n, lev.n))
    G = sp.csr_matrix((lev.val, lev.col, lev.rowptr), shape=(lev.n, lev.n))
    Gp = G[perm][:, perm].tocsr()  # producer matrix in the reference's numbering
    Gp.sort_indices()
    R.sort_indices()
    assert np.array_equal(R.indptr, Gp.indptr) and np.array_equal(R.indices, Gp.indices)
    scale = np.abs(R.data).max()
    # rounding of the reference's 27-point Gauss sums against exact products
    assert np.abs(R.data - Gp.data).max() <= 2e-12 * scale
    # Dirichlet rows: identity
    d = ref[p + "bnd1"]
    assert np.array_equal(np.sort(perm[d]), np.sort(lev.dirichlet))


@pytest.mark.parametrize("level,mesh_level", [(1, 3), (2, 2)])
def test_q2_patches_match_reference(ref, level, mesh_level):
    h = q2_hierarchy(3)
    lev = h.levels[level]
    p = f"L{mesh_level}_"
    pos = {k: i for i, k in enumerate(coord_keys(lev.coords))}
    perm = np.array([pos[k] for k in coord_keys(ref[p + "coords"])])
    mine = {tuple(row) for row in h.mesh.patches(level)}
    theirs = {tuple(perm[row]) for row in ref[p + "patches"]}
    assert mine == theirs  # same patches, same patch-local node order


def test_q2_reproduces_quadratics():
    """the triquadratic element reproduces quadratics exactly: on interior rows
    (A u)_i = int grad u . grad phi_i = -int (laplace u) phi_i, so for
    u = x^2 + y z (laplace 2) and v = x^2 / 2 (laplace 1), A u = 2 A v there;
    constants are in the kernel"""
    h = q2_hierarchy(3, colours=())
    lev = h.fine
    X = lev.coords
    A = sp.csr_matrix((lev.val, lev.col, lev.rowptr), shape=(lev.n, lev.n))
    assert np.abs(A @ np.ones(lev.n)).max() < 1e-12
    inner = np.all((X > 1e-9) & (X < 1 - 1e-9), axis=1)
    u = X[:, 0] ** 2 + X[:, 1] * X[:, 2]
    v = 0.5 * X[:, 0] ** 2
    Au, Av = A @ u, A @ v
    assert np.abs(Au[inner] - 2.0 * Av[inner]).max() < 1e-12
    # -A v = int phi_i, positive for the triquadratic Lagrange basis
    assert np.all(-Av[inner] > 0)


# ------------------------------------------------------------ adaptive levels
ADA = dict(adaptive=2, radius=0.25, centre=(0.5, 0.47, 0.53))


@pytest.fixture(scope="module")
def ada():
    return q2_hierarchy(3, **ADA)


def quadratic(X):
    x, y, z = X[:, 0], X[:, 1], X[:, 2]
    return x * x + y * z - 2 * x * z + 0.5 * y + 1.0


def test_adaptive_hanging_nodes_are_q2_exact(ada):
    """every constraint reproduces quadratics (the Q2 trace on the coarse face),
    no parent is itself constrained, and both adaptive levels have some"""
    for lev in ada.levels[-2:]:
        assert len(lev.hn_edges) > 0 and len(lev.hn_faces) > 0
        u = quadratic(lev.coords)
        w = np.array([0.375, 0.75, -0.125])
        e, f = lev.hn_edges, lev.hn_faces
        assert np.abs(u[e[:, 0]] - u[e[:, 1:]] @ w).max() < 1e-14
        wf = np.outer(w, w).reshape(-1)
        assert np.abs(u[f[:, 0]] - u[f[:, 1:]] @ wf).max() < 1e-14
        hanging = set(e[:, 0]) | set(f[:, 0])
        assert not (set(e[:, 1:].ravel()) | set(f[:, 1:].ravel())) & hanging
        # the constrained node sits a quarter of the way along its parents' edge
        # (midway between the near parent and the mid-edge parent), and the
        # face node at the matching point of its 3x3 parent patch
        X = lev.coords
        assert np.abs(X[e[:, 0]] - 0.5 * (X[e[:, 1]] + X[e[:, 2]])).max() < 1e-15
        assert np.abs(X[e[:, 3]] - 2.0 * X[e[:, 2]] + X[e[:, 1]]).max() < 1e-15
        assert np.abs(X[f[:, 0]] - np.einsum("k,nkd->nd", wf, X[f[:, 1:]])).max() < 1e-15


def test_adaptive_condensed_operator(ada):
    """the condensed operator equals T^T A T formed with scipy from the
    unconstrained patch assembly; hanging rows are identity rows; it is
    symmetric, has constants in its kernel and is exact on quadratics"""
    M = ada.mesh
    l = ada.nlevels - 1
    lev = ada.levels[l]
    n = lev.n
    rp, col, val = M.q2_condensed(l, ())
    Ac = sp.csr_matrix((val, col, rp), shape=(n, n))
    prp, pcol = M.patch_structure(l)
    # the unconstrained assembly: Dirichlet-free patch matrices
    A = sp.csr_matrix((M.assemble_q2(l, len(pcol), ()), pcol, prp), shape=(n, n))
    hang = np.zeros(n, bool)
    w = np.array([0.375, 0.75, -0.125])
    T = sp.lil_matrix((n, n))
    T.setdiag(1.0)
    for e in lev.hn_edges:
        hang[e[0]] = True
        T[e[0], e[0]] = 0.0
        for t in range(3):
            T[e[0], e[1 + t]] = w[t]
    wf = np.outer(w, w).reshape(-1)
    for f in lev.hn_faces:
        hang[f[0]] = True
        T[f[0], f[0]] = 0.0
        for t in range(9):
            T[f[0], f[1 + t]] = wf[t]
    T = T.tocsr()
    want = (T.T @ A @ T).tolil()
    for h in np.flatnonzero(hang):
        want[h, h] = 1.0
    want = want.tocsr()
    assert abs(Ac - want).max() <= 1e-12 * abs(want).max()
    H = np.flatnonzero(hang)
    assert np.all(np.diff(Ac.indptr)[H] == 1) and np.all(Ac[H, H].A1 == 1.0)
    assert abs(Ac - Ac.T).max() <= 1e-13 * abs(Ac).max()
    free = ~hang
    assert np.abs((Ac @ np.ones(n))[free]).max() < 1e-12
    X = lev.coords
    inner = free & np.all((X > 1e-9) & (X < 1 - 1e-9), axis=1)
    u, v = X[:, 0] ** 2 + X[:, 1] * X[:, 2], 0.5 * X[:, 0] ** 2
    assert np.abs((Ac @ u)[inner] - 2.0 * (Ac @ v)[inner]).max() < 1e-11


def test_adaptive_levels_are_nested(ada):
    """coarse nodes keep their numbers and positions; new nodes sit at the
    midpoints their transfer maps name"""
    for l in range(1, ada.nlevels):
        c2f, z, v, a = ada.levels[l].transfer
        Xc, Xf = ada.levels[l - 1].coords, ada.levels[l].coords
        assert np.array_equal(Xf[: len(Xc)], Xc) and np.array_equal(c2f, np.arange(len(Xc)))
        assert np.allclose(Xf[z[:, 0]], Xf[z[:, 1:]].mean(axis=1))
        assert np.allclose(Xf[v[:, 0]], Xf[v[:, 1:]].mean(axis=1))
        assert np.allclose(Xf[a[:, 0]], Xf[a[:, 1:]].mean(axis=1))
        assert len(c2f) + len(z) + len(v) + len(a) == len(Xf)


def test_adaptive_oracle_solve(ada):
    """the oracle's Newton/GMRES/MG with the hanging-node operations
    (HNStructureQ23d weights) converges on the adaptive hierarchy"""
    import oracle_ctypes as O
    from gascoignelib_amd.problem import lattice_rhs
    f = lattice_rhs(ada.fine.coords, 3)
    f[ada.fine.dirichlet] = 0.0
    om = O.MG(ada, mgtype="V", iterpre=2, iterpost=2, iterexact=10, omega=0.8, mgomega=1.0,
              ndirect=200)
    u = np.zeros_like(f)
    rc, rep = om.newton(u, f, O.cginfo(1e-10, 1e-30, 100), O.nlinfo(1e-10, 1e-30, 30), norm_kind=1)
    assert rc == 0 and rep.gmres_iterations < 40


def test_adaptive_with_every_patch_marked_is_uniform():
    """a ball that holds the whole domain marks every patch: the adaptive levels
    are then the uniform refinements, node for node and entry for entry, with
    no hanging nodes"""
    a = q2_hierarchy(2, adaptive=2, radius=10.0)
    u = q2_hierarchy(4)
    assert a.nlevels == u.nlevels
    for la, lu in zip(a.levels, u.levels):
        assert len(la.hn_edges) == 0 and len(la.hn_faces) == 0
        assert np.array_equal(la.coords, lu.coords)
        assert np.array_equal(la.rowptr, lu.rowptr) and np.array_equal(la.col, lu.col)
        assert np.abs(la.val - lu.val).max() <= 1e-14 * np.abs(lu.val).max()
        assert np.array_equal(np.sort(la.dirichlet), np.sort(lu.dirichlet))


def test_adaptive_step_that_marks_nothing_keeps_the_uniform_levels():
    """a ball outside the domain marks no node: the mesh is not refined, so the
    hierarchy is the uniform one (the reference keeps its levels)"""
    h = q2_hierarchy(2, adaptive=2, radius=0.25, centre=(5.0, 5.0, 5.0))
    u = q2_hierarchy(2)
    assert h.nlevels == u.nlevels
    for lh, lu in zip(h.levels, u.levels):
        assert np.array_equal(lh.coords, lu.coords) and np.array_equal(lh.col, lu.col)


@pytest.mark.parametrize("source", ["refine", "gup"])
def test_adaptive_levels_equal_reference(golden, tmp_path, source):
    """the adaptive hierarchy of the fixture's refinement (cube.inp, prerefine
    3, two refine_nodes steps of the nodes in the sphere of radius 0.08 around
    (0.41, 0.53, 0.47)) against the compiled reference's MeshAgent: made by
    the producer's own refinement, and read from the .gup file the
    reference's MeshAgent::write_gup wrote for that mesh"""
    import gzip
    import shutil
    import sys
    sys.path.insert(0, GOLDEN)
    import digests as D
    from gascoignelib_amd.structure import Mesh
    z = golden("q2_3d_8_ada")
    nl = int(z["nlevels"])
    if source == "refine":
        M = Mesh.adaptive("cube.inp", 3, 1, 2, 0.08, (0.41, 0.53, 0.47))
    else:
        gup = tmp_path / "mesh.gup"
        with gzip.open(os.path.join(GOLDEN, "q2_3d_8_ada.gup.gz"), "rb") as src, open(gup, "wb") as dst:
            shutil.copyfileobj(src, dst)
        M = Mesh.from_gup(gup, 1)
    assert M.nlevels == nl
    w1 = np.array([0.375, 0.75, -0.125])
    wf = np.outer(w1, w1).ravel()
    ranks_r, ranks_m = [], []
    for s in range(nl):
        m = nl - 1 - s
        xr, xm = z[f"L{m}_coords"], M.coords(s)
        rr, rm = D.rank_of(xr), D.rank_of(xm)
        ranks_r.append(rr)
        ranks_m.append(rm)
        assert len(xr) == len(xm)
        assert np.array_equal(xr[np.argsort(rr)], xm[np.argsort(rm)])
        cr = D.canonical_level(xr, cells=z[f"L{m}_cells"])[0]
        cm = D.canonical_level(xm, cells=M.cells(s))[0]
        assert cr["cells_sha256"] == cm["cells_sha256"]
        em, fm = M.hn(s)
        er, fr = z[f"L{m}_hn_edge"], z[f"L{m}_hn_face"]
        assert sorted(map(tuple, rr[er])) == sorted(map(tuple, rm[em]))
        pr = {rr[r[0]]: sorted(zip(rr[r[1:]], wf)) for r in fr}
        pm = {rm[r[0]]: sorted(zip(rm[r[1:]], wf)) for r in fm}
        assert pr == pm
    for s in range(1, nl):
        m = nl - 1 - s
        tr = D.canonical_transfer(*(z[f"L{m}_{k}"] for k in ("c2f", "zweier", "vierer", "achter")),
                                  ranks_r[s], ranks_r[s - 1])
        assert tr == D.canonical_transfer(*M.transfer(s), ranks_m[s], ranks_m[s - 1])


@pytest.mark.parametrize("kw", [dict(adaptive=3), dict(adaptive=2, radius=-1.0)])
def test_adaptive_bad_arguments_raise(kw):
    with pytest.raises(RuntimeError, match="ghs_build_adaptive"):
        q2_hierarchy(2, **kw)


def test_gup_errors(tmp_path):
    from gascoignelib_amd.structure import Mesh
    bad = tmp_path / "bad.gup"
    bad.write_text("2 dimension\n")
    with pytest.raises(RuntimeError, match="not a 3D .gup"):
        Mesh.from_gup(bad)
    with pytest.raises(RuntimeError, match="cannot open"):
        Mesh.from_gup(tmp_path / "missing.gup")
