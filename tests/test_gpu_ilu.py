"""ILU(0) (SparseBlockIlu, src/LinAlg/sparseblockilu.xx) on the GPU.

* On Gascoigne's own meshes (compiled-reference fixtures: Q1 2D, Q1 3D, the
  3x3x3-hex cube, Q1 and Q2 with hanging nodes), in the identity and the
  reversed order: the permuted structure, the factors after compute_ilu and one
  solve equal the reference's bit for bit.
* Blocks of 2 and 4 components in a random order: factors and FP64 solves
  equal the oracle's bit for bit; FP32 solves within 1e-5.
* The level schedules are the longest paths of the lower and upper
  dependency graphs (computed here by brute force), and a tridiagonal matrix
  (one chain) has n levels.
* In the colour order (ghs_colour_order) a uniform level's triangles have 8
  levels each, and the factors and solves still equal the oracle's.
* The V-cycle with the ILU smoother (StdSolver::smooth "ilu",
  src/Solver/stdsolver.cc:962-967) equals the oracle's bit for bit, in the
  caller's and in the lexicographic device numbering, for the identity, the
  reversed and the colour orders; Newton / GMRES iterate with the oracle's
  counts; the reduced-precision solver converges.
* A zero pivot and an order that is not a permutation are errors.
"""
import ctypes as C

import numpy as np
import pytest

import oracle_ctypes as O
from gascoignelib_amd import hip as H
from gascoignelib_amd import native as N
from gascoignelib_amd.problem import Hierarchy, Level, block_operator
from gascoignelib_amd.structure import Mesh, assemble_hierarchy, colour_order, uniform_hierarchy

pytestmark = pytest.mark.gpu

CASES = ["q1_2d_128", "q1_3d_16", "q1_3d_cube3_12", "q1_3d_8_ada", "q2_3d_8_ada"]
MGP = dict(mgtype="V", iterpre=2, iterpost=2, iterexact=4, omega=0.8, mgomega=1.0)


@pytest.fixture(scope="module")
def ctx():
    c = H.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("name", CASES)
def test_ilu_equals_reference(ctx, golden, name):
    print(f"[ilu reference {name}]", flush=True)
    z = golden(name)
    checked = 0
    for l in range(int(z["nlevels"])):
        p = f"L{l}_"
        if p + "ilu0_val" not in z.files:
            continue
        rp, col = z[p + "rowptr"], z[p + "col"]
        A = H.CudaCSRMatrixInterface(ctx, rp, col, z[p + "val"])
        n = len(rp) - 1
        for kind in (0, 1):
            q = f"{p}ilu{kind}_"
            perm = np.arange(n) if kind == 0 else np.arange(n)[::-1]
            I = H.SparseBlockIlu(ctx, A, perm)
            frp, fcol, fdg, fval = I.factors(len(col))
            assert np.array_equal(frp, z[q + "rowptr"])
            assert np.array_equal(fcol, z[q + "col"])
            assert np.array_equal(fdg, z[q + "diag"])
            assert np.array_equal(fval, z[q + "val"])
            X = H.CudaVectorInterface.from_global(ctx, z[p + "ilu_in"])
            I.solve(X)
            assert np.array_equal(X.copy_back(), z[q + "solve"])
            checked += 1
    assert checked >= 2


def longest_paths(rp, col, dg):
    """level of each row in the lower (forward) and upper (backward) graphs"""
    n = len(rp) - 1
    lf = np.zeros(n, np.int64)
    for i in range(n):
        lo = col[rp[i]:dg[i]]
        lf[i] = 1 + lf[lo].max() if len(lo) else 0
    lb = np.zeros(n, np.int64)
    for i in range(n - 1, -1, -1):
        up = col[dg[i] + 1:rp[i + 1]]
        lb[i] = 1 + lb[up].max() if len(up) else 0
    return int(lf.max()) + 1, int(lb.max()) + 1


def cube_level(pre, B, seed):
    M = Mesh("cube.inp", 3, pre, 1)
    h = assemble_hierarchy(M)
    lev = h.fine
    if B == 1:
        return lev, lev.val
    rng = np.random.default_rng(seed)
    G = rng.uniform(-1, 1, (B, B))
    Cm, S = G @ G.T + B * np.eye(B), rng.uniform(-0.3, 0.3, (B, B))
    return lev, block_operator(lev, Cm, S)


@pytest.mark.parametrize("B", [1, 2, 4])
def test_block_ilu_random_order(ctx, B):
    print(f"[ilu blocks B={B}]", flush=True)
    lev, val = cube_level(4 if B == 1 else 3, B, B)
    n = lev.n
    perm = np.random.default_rng(3 + B).permutation(n).astype(np.int32)
    A = H.CudaCSRMatrixInterface(ctx, lev.rowptr, lev.col, val, B)
    I = H.SparseBlockIlu(ctx, A, perm)
    Io = O.Ilu(O.Matrix(lev.rowptr, lev.col, val, B), perm)
    rp, col, dg, fv = I.factors(len(lev.col))
    assert np.array_equal(rp, Io.rowptr) and np.array_equal(col, Io.col)
    assert np.array_equal(dg, Io.diag)
    assert np.array_equal(fv, Io.val)
    assert I.levels() == longest_paths(rp, col, dg)
    x = np.random.default_rng(7).uniform(-1, 1, n * B)
    X = H.CudaVectorInterface.from_global(ctx, x, B)
    I.solve(X)
    xo = x.copy()
    Io.solve(xo)
    assert np.array_equal(X.copy_back(), xo)
    X32 = H.CudaVectorInterface.from_global(ctx, x, B, "f32")
    I.solve(X32)
    assert np.abs(X32.copy_back() - xo).max() <= 1e-5 * np.abs(xo).max()


def test_colour_order_levels_and_solve(ctx):
    """in the colour order each triangle has one level per colour; factors and
    solves still equal the oracle's, and the FP32 solve is close"""
    print("[ilu colour order]", flush=True)
    lev = uniform_hierarchy("cube.inp", 3, 5, drop_coarse=1).fine
    perm, nc = colour_order(lev.rowptr, lev.col)
    A = H.CudaCSRMatrixInterface(ctx, lev.rowptr, lev.col, lev.val)
    I = H.SparseBlockIlu(ctx, A, perm)
    assert I.levels() == (nc, nc) == (8, 8)
    Io = O.Ilu(O.Matrix(lev.rowptr, lev.col, lev.val), perm)
    assert np.array_equal(I.factors(len(lev.col))[3], Io.val)
    x = np.random.default_rng(5).uniform(-1, 1, lev.n)
    X = H.CudaVectorInterface.from_global(ctx, x)
    I.solve(X)
    xo = x.copy()
    Io.solve(xo)
    assert np.array_equal(X.copy_back(), xo)
    X32 = H.CudaVectorInterface.from_global(ctx, x, 1, "f32")
    I.solve(X32)
    assert np.abs(X32.copy_back() - xo).max() <= 1e-5 * np.abs(xo).max()


def test_tridiagonal_chain(ctx):
    """one dependency chain: n levels each way, and ILU(0) is the exact LU"""
    n = 300
    rng = np.random.default_rng(1)
    d = 4.0 + rng.uniform(0, 1, n)
    e = rng.uniform(-1, 1, n - 1)
    rowptr, col, val = [0], [], []
    for i in range(n):
        for j, v in ((i - 1, e[i - 1] if i else 0), (i, d[i]), (i + 1, e[i] if i < n - 1 else 0)):
            if 0 <= j < n:
                col.append(j)
                val.append(v)
        rowptr.append(len(col))
    A = H.CudaCSRMatrixInterface(ctx, rowptr, col, val)
    I = H.SparseBlockIlu(ctx, A)
    assert I.levels() == (n, n)
    b = rng.uniform(-1, 1, n)
    X = H.CudaVectorInterface.from_global(ctx, b)
    I.solve(X)
    import scipy.sparse as sp
    Ad = sp.csr_matrix((val, col, rowptr), shape=(n, n))
    np.testing.assert_allclose(Ad @ X.copy_back(), b, rtol=0, atol=1e-13)


def test_ilu_errors(ctx):
    print("[ilu errors]", flush=True)
    lib = N.hip()
    lev, val = cube_level(2, 1, 0)
    A = H.CudaCSRMatrixInterface(ctx, lev.rowptr, lev.col, val)
    h = C.c_void_p()
    bad = np.zeros(lev.n, np.int32)  # not a permutation
    assert lib.gh_ilu_create(ctx.h, A.h, N.ptr(bad, N.i32p), C.byref(h)) == N.GH_EINVAL
    # a zero diagonal in the first row (a host-only matrix: the device matrix
    # refuses it at creation)
    v0 = val.copy()
    for q in range(lev.rowptr[0], lev.rowptr[1]):
        if lev.col[q] == 0:
            v0[q] = 0.0
    Z = H.CudaCSRMatrixInterface.host_only(ctx, lev.rowptr, lev.col, v0)
    rc = lib.gh_ilu_create(ctx.h, Z.h, None, C.byref(h))
    assert rc == N.GH_ESINGULAR, lib.gh_last_error().decode()
    # [[1, 1], [1, 1]]: nonzero diagonal, the second pivot 1 - 1 = 0
    Z2 = H.CudaCSRMatrixInterface(ctx, [0, 2, 4], [0, 1, 0, 1], [1.0, 1.0, 1.0, 1.0])
    rc = lib.gh_ilu_create(ctx.h, Z2.h, None, C.byref(h))
    assert rc == N.GH_ESINGULAR, lib.gh_last_error().decode()
    # vector of the wrong shape
    I = H.SparseBlockIlu(ctx, A)
    X = H.CudaVectorInterface(ctx, lev.n + 1)
    assert lib.gh_ilu_solve(ctx.h, I.h, X.h) == N.GH_EINVAL


def mg_hierarchy(pre, B):
    M = Mesh("cube.inp", 3, pre, 1)
    h = assemble_hierarchy(M)
    if B == 1:
        return h
    rng = np.random.default_rng(B)
    G = rng.uniform(-1, 1, (B, B))
    Cm, S = G @ G.T + B * np.eye(B), rng.uniform(-0.3, 0.3, (B, B))
    levels = [Level(lev.n, B, lev.rowptr, lev.col, block_operator(lev, Cm, S), lev.transfer,
                    lev.hn_kind, dirichlet=lev.dirichlet,
                    dirichlet_comps=np.arange(B, dtype=np.int32), coords=lev.coords)
              for lev in h.levels]
    return Hierarchy(levels, 3, f"cube{pre}:B{B}")


@pytest.mark.parametrize("B", [1, 4])
@pytest.mark.parametrize("order", [None, "lexicographic"])
@pytest.mark.parametrize("ilu_order", ["identity", "reversed", "colour"])
def test_vcycle_with_ilu_vs_oracle(ctx, B, order, ilu_order):
    print(f"[ilu V-cycle B={B} order={order} ilu={ilu_order}]", flush=True)
    h = mg_hierarchy(4 if B == 1 else 3, B)
    orders = None
    if ilu_order == "reversed":
        orders = [np.arange(h.level(l).n, dtype=np.int32)[::-1].copy() for l in range(h.nlevels)]
    elif ilu_order == "colour":
        orders = [colour_order(h.level(l).rowptr, h.level(l).col)[0] for l in range(h.nlevels)]
    f = np.random.default_rng(9).uniform(-1, 1, h.fine.n * B)
    for c in range(B):
        f[h.fine.dirichlet * B + c] = 0.0
    om = O.MG(h, ndirect=1, smoother="ilu", ilu_orders=orders, **MGP)
    xo = np.zeros_like(f)
    om.precondition(xo, f)
    mg = H.CudaMultiLevelSolver(ctx, h, ndirect=1, linear_smooth="ilu", ilu_orders=orders,
                                order=order, **MGP)
    X = mg.new_vector()
    mg.precondition(X, H.CudaVectorInterface.from_global(ctx, f, B))
    assert np.array_equal(X.copy_back(), xo)
    uo = np.zeros_like(f)
    _, orep = om.newton(uo, f, O.cginfo(1e-10, 1e-30, 100), O.nlinfo(1e-10, 1e-30, 30),
                        norm_kind=1)
    U = mg.new_vector()
    rep = mg.Solve(U, H.CudaVectorInterface.from_global(ctx, f, B), H.CGInfo(1e-10, 1e-30, 100),
                   H.NLInfo(1e-10, 1e-30, 30), norm_kind=1)
    assert rep.newton_iterations == orep.newton_iterations
    assert abs(rep.gmres_iterations - orep.gmres_iterations) <= 1
    assert np.linalg.norm(U.copy_back() - uo) <= 1e-12 * np.linalg.norm(uo)


def test_ilu_mixed_precision_solve(ctx):
    print("[ilu mixed]", flush=True)
    h = mg_hierarchy(4, 1)
    f = np.random.default_rng(11).uniform(-1, 1, h.fine.n)
    f[h.fine.dirichlet] = 0.0
    om = O.MG(h, ndirect=1, smoother="ilu", **MGP)
    uo = np.zeros_like(f)
    om.newton(uo, f, O.cginfo(1e-12, 1e-30, 100), O.nlinfo(1e-12, 1e-30, 30), norm_kind=1)
    mg = H.CudaMultiLevelSolver(ctx, h, ndirect=1, linear_smooth="ilu", prec="f32",
                                krylov="fgmres", **MGP)
    U = mg.new_vector()
    rep = mg.Solve(U, H.CudaVectorInterface.from_global(ctx, f), H.CGInfo(0.9e-10, 1e-30, 100),
                   H.NLInfo(1e-10, 1e-30, 30), norm_kind=1)
    assert H.N.NL_STATUS[rep.status] == "converged"
    assert np.linalg.norm(U.copy_back() - uo) <= 1e-8 * np.linalg.norm(uo)
