"""Device assembly of triquadratic patch elements (gh_assemble_q2, gh_form_q2)
on the GPU.

The reference assembles Laplace<3> with CGDiscQ23d (src/DofHandler/cgdisc.h:
1292; CGDisc::Matrix, cgdisc.h:599-635, with SparseBlockMatrix::dirichlet,
src/LinAlg/sparseblockmatrix.xx:205-218) on the CPU.  Checks:

* the device values equal the host producer's Q2 assembly bit for bit, Dirichlet
  rows and columns included (the host producer equals the reference's Q23d
  output to rounding, tests/test_q2_structure.py);
* with one coefficient per patch element, bit for bit against a numpy
  restatement that adds k_p E^p_ab to each entry in ascending patch order,
  in FP64 and in f16 storage; a coefficient of 2 doubles every free entry
  exactly;
* the element-by-element residual f - A(u) against the assembled residual;
* misuse is an error: a pattern that lacks a coupling, a coefficient vector
  of the wrong size, 2D patches.
"""
import numpy as np
import pytest
import scipy.sparse as sp

from gascoignelib_amd import hip as H
from gascoignelib_amd import native as N
from gascoignelib_amd.structure import Mesh

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = H.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def level():
    M = Mesh("cube.inp", 3, 4, 1)
    l = M.nlevels - 1
    rp, col = M.patch_structure(l)
    host = M.assemble_q2(l, len(col), (1,))
    free = M.assemble_q2(l, len(col), ())
    return dict(M=M, coords=M.coords(l), P=M.patches(l), rp=rp, col=col, host=host, free=free,
                d=M.dirichlet(l, (1,)), n=len(rp) - 1)


def element_matrix(h):
    """E_ab of the patch element with steps h (x, y, z), the device's operations"""
    s = np.array([[7.0, -8.0, 1.0], [-8.0, 16.0, -8.0], [1.0, -8.0, 7.0]])
    m = np.array([[4.0, 2.0, -1.0], [2.0, 16.0, 2.0], [-1.0, 2.0, 4.0]])
    S, Mm = [], []
    for hd in h:
        L = 2.0 * hd
        S.append(s / (3.0 * L))
        Mm.append((m * L) / 30.0)
    a = np.arange(27)
    ax, ay, az = a % 3, (a // 3) % 3, a // 9
    X = lambda T, d, i: T[d][i[:, None], i[None, :]]  # noqa: E731
    t1 = X(S, 0, ax) * X(Mm, 1, ay) * X(Mm, 2, az)
    t2 = X(Mm, 0, ax) * X(S, 1, ay) * X(Mm, 2, az)
    t3 = X(Mm, 0, ax) * X(Mm, 1, ay) * X(S, 2, az)
    return (t1 + t2) + t3


def numpy_assembly(L, k):
    """sum over patches in ascending order of k_p E^p, then the Dirichlet rows
    and columns (np.add.at adds in array order: patch by patch)"""
    P, x, rp, col, n = L["P"], L["coords"], L["rp"], L["col"], L["n"]
    h = np.stack([x[P[:, 1], 0] - x[P[:, 0], 0], x[P[:, 3], 1] - x[P[:, 0], 1],
                  x[P[:, 9], 2] - x[P[:, 0], 2]], axis=1)
    rows = np.repeat(np.arange(n), np.diff(rp)).astype(np.int64)
    keys = rows * n + col
    val = np.zeros(len(col))
    shapes = {}
    for p in range(len(P)):
        key = tuple(h[p])
        if key not in shapes:
            shapes[key] = element_matrix(h[p])
        E = shapes[key]
        q = np.searchsorted(keys, P[p][:, None].astype(np.int64) * n + P[p][None, :])
        np.add.at(val, q.ravel(), (k[p] * E).ravel())
    isd = np.zeros(n, bool)
    isd[L["d"]] = True
    val[isd[col]] = 0.0
    val[isd[rows]] = 0.0
    val[isd[rows] & (rows == col)] = 1.0
    return val


def test_q2_assembly_equals_host_producer(ctx, level):
    print("[device assembly q2]", flush=True)
    L = level
    A = H.CudaCSRMatrixInterface(ctx, L["rp"], L["col"], np.where(L["host"] != 0.0, L["host"], 1.0))
    Pt = H.DwrPatches(ctx, 3, L["coords"], L["P"])
    Pt.assemble_q2(A, mask=H.DirichletMask(ctx, L["n"], 1, L["d"]))
    assert np.array_equal(A.values(), L["host"])
    # the refreshed layout runs the product
    x = np.random.default_rng(3).uniform(-1, 1, L["n"])
    X, Y = H.CudaVectorInterface.from_global(ctx, x), H.CudaVectorInterface(ctx, L["n"])
    A.vmult(Y, X)
    ref = sp.csr_matrix((L["host"], L["col"], L["rp"]), shape=(L["n"], L["n"])) @ x
    np.testing.assert_allclose(Y.copy_back(), ref, rtol=0, atol=1e-12 * np.abs(ref).max())


@pytest.mark.parametrize("prec", ["f64", "f16"])
def test_q2_assembly_per_patch_coefficient_bitexact(ctx, level, prec):
    print(f"[device assembly q2 k {prec}]", flush=True)
    L = level
    k = np.exp(np.random.default_rng(5).normal(size=len(L["P"])))
    want = numpy_assembly(L, k)
    A = H.CudaCSRMatrixInterface(ctx, L["rp"], L["col"], np.where(L["host"] != 0.0, L["host"], 1.0),
                                 prec=prec)
    Pt = H.DwrPatches(ctx, 3, L["coords"], L["P"])
    Pt.assemble_q2(A, k=H.CudaVectorInterface.from_global(ctx, k),
                   mask=H.DirichletMask(ctx, L["n"], 1, L["d"]))
    got = A.values()
    if prec == "f64":
        assert np.array_equal(got, want)
    else:  # stored as the same values uploaded in f16
        Ref = H.CudaCSRMatrixInterface(ctx, L["rp"], L["col"], want, prec=prec)
        assert np.array_equal(got, Ref.values())
    # k = 2: every free entry doubles exactly, Dirichlet rows stay unit rows
    A2 = H.CudaCSRMatrixInterface(ctx, L["rp"], L["col"], np.where(L["host"] != 0.0, L["host"], 1.0))
    Pt.assemble_q2(A2, k=H.CudaVectorInterface.from_global(ctx, np.full(len(L["P"]), 2.0)),
                   mask=H.DirichletMask(ctx, L["n"], 1, L["d"]))
    rows = np.repeat(np.arange(L["n"]), np.diff(L["rp"]))
    isd = np.zeros(L["n"], bool)
    isd[L["d"]] = True
    v2, v1 = A2.values(), L["host"]
    assert np.array_equal(v2[~isd[rows]], 2.0 * v1[~isd[rows]])
    assert np.array_equal(v2[isd[rows]], v1[isd[rows]])


def test_q2_form_matches_assembled_residual(ctx, level):
    print("[form q2]", flush=True)
    L = level
    n = L["n"]
    rng = np.random.default_rng(7)
    k = np.exp(rng.normal(size=len(L["P"])))
    u = rng.uniform(-1, 1, n)
    f = rng.uniform(-1, 1, n)
    Pt = H.DwrPatches(ctx, 3, L["coords"], L["P"])
    K = H.CudaVectorInterface.from_global(ctx, k)
    U, F, R = (H.CudaVectorInterface.from_global(ctx, v) for v in (u, f, np.zeros(n)))
    Pt.form_q2(R, U, F, k=K, mask=H.DirichletMask(ctx, n, 1, L["d"]))
    # the unconstrained operator with the same coefficients
    Lf = dict(L, d=np.zeros(0, np.int32))
    A = sp.csr_matrix((numpy_assembly(Lf, k), L["col"], L["rp"]), shape=(n, n))
    r = f - A @ u
    r[L["d"]] = 0.0
    got = R.copy_back()
    np.testing.assert_allclose(got, r, rtol=0, atol=1e-13 * np.abs(A @ u).max())
    assert np.all(got[L["d"]] == 0.0)


def test_q2_assembly_errors(ctx, level):
    print("[device assembly q2 errors]", flush=True)
    lib = N.hip()
    L = level
    n = L["n"]
    Pt = H.DwrPatches(ctx, 3, L["coords"], L["P"])
    # a diagonal-only pattern misses the couplings of the elements
    Dg = H.CudaCSRMatrixInterface(ctx, np.arange(n + 1), np.arange(n), np.ones(n))
    rc = lib.gh_assemble_q2(ctx.h, Pt.h, None, None, Dg.h)
    assert rc == N.GH_EINVAL and "pattern" in lib.gh_last_error().decode()
    # a coefficient vector of the wrong size
    A = H.CudaCSRMatrixInterface(ctx, L["rp"], L["col"], np.where(L["host"] != 0.0, L["host"], 1.0))
    K = H.CudaVectorInterface(ctx, len(L["P"]) + 1)
    assert lib.gh_assemble_q2(ctx.h, Pt.h, K.h, None, A.h) == N.GH_EINVAL
    # 2D patches are refused (the patch element is triquadratic)
    M2 = Mesh("square.inp", 2, 3, 1)
    l2 = M2.nlevels - 1
    P2 = H.DwrPatches(ctx, 2, M2.coords(l2), M2.patches(l2))
    rp2, col2 = M2.structure(l2)
    A2 = H.CudaCSRMatrixInterface(ctx, rp2, col2, np.ones(len(col2)))
    assert lib.gh_assemble_q2(ctx.h, P2.h, None, None, A2.h) == N.GH_EINVAL
    # the refused calls left the matrix usable
    Pt.assemble_q2(A, mask=H.DirichletMask(ctx, n, 1, L["d"]))
    assert np.array_equal(A.values(), L["host"])


def q2_levels_with_k(M, ks, colours=(1,)):
    """a uniform Q2 hierarchy whose level operators carry per-patch
    coefficients ks[l] (numpy restatement of the device assembly)"""
    from gascoignelib_amd.problem import Hierarchy, Level
    levels = []
    for l in range(M.nlevels):
        rp, col = M.patch_structure(l)
        L = dict(P=M.patches(l), coords=M.coords(l), rp=rp, col=col, n=len(rp) - 1,
                 d=M.dirichlet(l, colours))
        val = numpy_assembly(L, ks[l])
        levels.append(Level(n=L["n"], ncomp=1, rowptr=rp, col=col, val=val,
                            transfer=M.transfer(l) if l > 0 else None, hn_kind=23,
                            hn_edges=np.zeros((0, 4), np.int32),
                            hn_faces=np.zeros((0, 10), np.int32), dirichlet=L["d"],
                            dirichlet_comps=np.zeros(1, np.int32), coords=L["coords"]))
    h = Hierarchy(levels, 3, "q2k")
    h.mesh = M
    return h


@pytest.mark.parametrize("prec,krylov", [("f64", "gmres"), ("f16", "fgmres")])
def test_q2_solver_reassembly_matches_host_built_solver(ctx, prec, krylov):
    print(f"[mg re-assembly q2 {prec} {krylov}]", flush=True)
    M = Mesh("cube.inp", 3, 4, 1)
    rng = np.random.default_rng(9)
    k1 = [np.exp(rng.normal(size=len(M.patches(l)))) for l in range(M.nlevels)]
    k2 = [np.exp(0.5 * rng.normal(size=len(M.patches(l)))) for l in range(M.nlevels)]
    h1, h2 = q2_levels_with_k(M, k1), q2_levels_with_k(M, k2)
    params = dict(mgtype="V", iterpre=2, iterpost=2, iterexact=10, omega=0.8, mgomega=1.0,
                  ndirect=200, prec=prec, krylov=krylov, order="lexicographic")
    mg1 = H.CudaMultiLevelSolver(ctx, h1, **params)
    mg2 = H.CudaMultiLevelSolver(ctx, h2, **params)
    pts = [H.DwrPatches(ctx, 3, M.coords(l), M.patches(l)) for l in range(M.nlevels)]
    mg1.assemble_q2(pts, [H.CudaVectorInterface.from_global(ctx, k) for k in k2])
    f = np.random.default_rng(4).uniform(-1, 1, h2.fine.n)
    f[h2.fine.dirichlet] = 0.0
    F = H.CudaVectorInterface.from_global(ctx, f)
    sols, reps = [], []
    for mg in (mg1, mg2):
        U = mg.new_vector()
        rep = mg.Solve(U, F, H.CGInfo(1e-10, 1e-30, 60), H.NLInfo(1e-10, 1e-30, 20),
                       vmax=10 if krylov == "gmres" else 20, norm_kind=1)
        sols.append(U.copy_back())
        reps.append((rep.newton_iterations, rep.gmres_iterations, H.N.NL_STATUS[rep.status]))
    assert reps[0] == reps[1] and reps[0][2] == "converged"
    assert np.array_equal(sols[0], sols[1])
    for l in range(h1.nlevels):
        for fp64 in (True, False):
            assert np.array_equal(mg1.level_matrix(l, fp64).values(),
                                  mg2.level_matrix(l, fp64).values())


def test_q2_condensed_over_hanging_nodes_equals_host(ctx):
    """gh_assemble_q2 (unconstrained) then gh_condense_q2 on every level of an
    adaptive Q2 hierarchy (8^3 + 2 refinements in a ball): the condensed
    values equal the host producer's (M.q2_condensed) bit for bit, the sign
    of an exactly-zero sum aside"""
    print("[condense q2]", flush=True)
    lib = N.hip()
    M = Mesh.adaptive("cube.inp", 3, 1, 2, 0.25, (0.6, 0.45, 0.4))
    checked = 0
    for l in range(M.nlevels):
        e, f = M.hn(l)
        rp, col = M.patch_structure(l)
        free = M.assemble_q2(l, len(col), ())
        A = H.CudaCSRMatrixInterface(ctx, rp, col, np.where(free != 0.0, free, 1.0))
        Pt = H.DwrPatches(ctx, 3, M.coords(l), M.patches(l))
        Pt.assemble_q2(A)
        assert np.array_equal(A.values(), free)
        rpc, colc, want = M.q2_condensed(l, (1,))
        n = len(rpc) - 1
        Ac = H.CudaCSRMatrixInterface(ctx, rpc, colc, np.where(want != 0.0, want, 1.0))
        mask = H.DirichletMask(ctx, n, 1, M.dirichlet(l, (1,)))
        H.condense_q2(ctx, e, f, A, Ac, mask)
        assert np.array_equal(Ac.values(), want), f"level {l}"
        checked += len(e) + len(f) > 0
    assert checked >= 1  # levels with hanging nodes were among them
    # a constraint naming a node outside the matrix is refused
    bad = np.array([[len(rpc) + 5, 0, 1, 2]], dtype=np.int32)
    rc = lib.gh_condense_q2(ctx.h, 1, N.ptr(bad, N.i32p), 0, None, A.h, None, Ac.h)
    assert rc == N.GH_EINVAL
    # and so is a node that two constraints name
    twice = np.array([[3, 0, 1, 2], [3, 4, 5, 6]], dtype=np.int32)
    rc = lib.gh_condense_q2(ctx.h, 2, N.ptr(twice, N.i32p), 0, None, A.h, None, Ac.h)
    assert rc == N.GH_EINVAL
