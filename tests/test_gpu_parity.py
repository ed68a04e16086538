"""HIP path vs the oracle and the compiled reference's golden vectors (GPU).

Tolerances (the bar of the FP64 path):
  * integer structure, grid transfer and hanging-node operators: bit-exact
    (their weights are powers of two or evaluated without contraction, and each
    output is summed in the reference's order);
  * FP64 SpMV / residual / Jacobi: bit-exact (one lane per row sums in the
    stored column order without contraction, as the CPU loop does);
  * the FP64 V-cycle with an iterative coarse solve: bit-exact; with the dense
    direct coarse solve and through GMRES (parallel dot products): 1e-12;
  * V-cycle, GMRES and Newton iterates: 1e-12 relative to the oracle,
    iteration counts within one.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle_ctypes as O
from conftest import fixture_path
from gascoignelib_amd import hip as H
from gascoignelib_amd.problem import Hierarchy, splitmix64_uniform

pytestmark = pytest.mark.gpu

FIXTURES = ["q1_2d_128", "q1_3d_16", "q1_3d_8_ada", "q2_3d_8_ada", "q1_3d_cube3_12",
            "q1_2d_16_ada", "q2_2d_16_ada"]
EPS = np.finfo(np.float64).eps


@pytest.fixture(scope="module")
def ctx():
    c = H.Context(0)
    yield c
    c.close()


def vec(ctx, a, prec="f64"):
    return H.CudaVectorInterface.from_global(ctx, a, 1, prec)


def row_bound(z, p, x, b):
    A = sp.csr_matrix((np.abs(z[p + "val"]), z[p + "col"], z[p + "rowptr"]))
    m = np.diff(z[p + "rowptr"])
    return (m + 2) * EPS * (A @ np.abs(x) + np.abs(b))


@pytest.mark.parametrize("name", FIXTURES)
def test_spmv_residual_jacobi_vs_reference(ctx, golden, name):
    z = golden(name)
    for l in range(int(z["nlevels"])):
        p = f"L{l}_"
        print(f"[{name} level {l}]", flush=True)
        A = H.CudaCSRMatrixInterface(ctx, z[p + "rowptr"], z[p + "col"], z[p + "val"])
        x, b = z[p + "x"], z[p + "b"]
        bound = row_bound(z, p, x, b)
        X, B = vec(ctx, x), vec(ctx, b)
        Y = H.CudaVectorInterface(ctx, len(x))
        A.vmult(Y, X, 1.0, 0.0)
        assert np.array_equal(Y.copy_back(), z[p + "Ax"])
        R = H.CudaVectorInterface(ctx, len(x))
        A.residual(R, X, B)
        r = R.copy_back()
        assert np.array_equal(r, z[p + "res"])
        # the smoother's fused sweep x' = x + omega D^-1 (b - A x), from its parts
        Xn = H.CudaVectorInterface(ctx, len(x))
        Xn.copy_from(X)
        Hh = H.CudaVectorInterface(ctx, len(x))
        A.smooth(Xn, B, Hh, 1, 0.8)
        assert np.array_equal(Xn.copy_back(), x + 0.8 * z[p + "jac"])
        # FP32 path: within the summation bound of FP32
        Af = H.CudaCSRMatrixInterface(ctx, z[p + "rowptr"], z[p + "col"], z[p + "val"], 1, "f32")
        Xf, Yf = vec(ctx, x, "f32"), H.CudaVectorInterface(ctx, len(x), 1, "f32")
        Af.vmult(Yf, Xf, 1.0, 0.0)
        assert np.all(np.abs(Yf.copy_back() - z[p + "Ax"]) <= 2e9 * bound)  # eps32/eps64 ~ 5e8
        # Jacobi on the reference's own residual is elementwise: exact
        R.assign(z[p + "res"])
        A.jacobi(R)
        assert np.array_equal(R.copy_back(), z[p + "jac"])
        # y = d*A*x + b*y
        Y.assign(b)
        A.vmult(Y, X, -2.0, 0.5)
        assert np.all(np.abs(Y.copy_back() - (-2.0 * z[p + "Ax"] + 0.5 * b)) <= 4 * bound)


@pytest.mark.parametrize("name", FIXTURES)
def test_transfer_bitexact_vs_reference(ctx, golden, name):
    z = golden(name)
    for l in range(int(z["nlevels"]) - 1):
        p, q = f"L{l}_", f"L{l + 1}_"
        nc = len(z[q + "x"])
        T = H.CudaMgInterpolatorNested(ctx, nc, z[p + "c2f"], z[p + "zweier"], z[p + "vierer"],
                                       z[p + "achter"])
        UF = vec(ctx, z[p + "t_in"])
        UC = H.CudaVectorInterface(ctx, nc)
        T.restrict_zero(UC, UF)
        assert np.array_equal(UC.copy_back(), z[p + "restrict"])
        UC.assign(z[q + "t_in"])
        UF.zero()
        T.prolongate_add(UF, UC)
        assert np.array_equal(UF.copy_back(), z[p + "prolongate"])


@pytest.mark.parametrize("name,kind", [("q1_3d_8_ada", 13), ("q2_3d_8_ada", 23),
                                       ("q1_2d_16_ada", 12), ("q2_2d_16_ada", 22)])
def test_hanging_nodes_bitexact_vs_reference(ctx, golden, name, kind):
    z = golden(name)
    for l in range(int(z["nlevels"])):
        p = f"L{l}_"
        if p + "hn_average" not in z.files:
            continue
        n = len(z[p + "hn_in"])
        Hs = H.HNStructure(ctx, kind, n, z[p + "hn_edge"], z[p + "hn_face"])
        for op in ("average", "distribute", "zero"):
            U = vec(ctx, z[p + "hn_in"])
            getattr(Hs, op)(U)
            assert np.array_equal(U.copy_back(), z[p + "hn_" + op]), op


def test_vector_ops(ctx):
    rng = np.random.default_rng(1)
    x, y = rng.standard_normal(100003), rng.standard_normal(100003)
    X, Y = vec(ctx, x), vec(ctx, y)
    assert X.norm() == pytest.approx(np.linalg.norm(x), rel=1e-14)
    assert X.dot(Y) == pytest.approx(x @ y, rel=1e-12)
    assert X.norm_l8() == np.abs(x).max()
    X.add(0.5, Y)
    assert np.allclose(X.copy_back(), x + 0.5 * y, rtol=1e-15)
    X.sadd(-1.0, 1.0, Y)
    assert np.allclose(X.copy_back(), -(x + 0.5 * y) + y, rtol=1e-15, atol=1e-15)
    X.scal(3.0)
    X.equ(2.0, Y)
    assert np.array_equal(X.copy_back(), 2.0 * y)
    Z = H.CudaVectorInterface(ctx, len(x), 1, "f32")
    Z.copy_from(Y)
    assert np.array_equal(Z.copy_back(), y.astype(np.float32).astype(np.float64))
    # the reductions are reproducible bit for bit
    assert X.dot(Y) == X.dot(Y)


def test_dirichlet_mask(ctx):
    x = np.arange(20, dtype=np.float64) + 1.0
    M = H.DirichletMask(ctx, 10, 2, [1, 4, 7], [1])
    X = vec(ctx, x)
    X2 = H.CudaVectorInterface.from_global(ctx, x, 2)
    M.zero(X2)
    e = x.copy()
    e[[3, 9, 15]] = 0.0
    assert np.array_equal(X2.copy_back(), e)
    del X


def test_spmv_transpose(ctx, golden):
    z = golden("q1_3d_8_ada")
    p = "L0_"
    A = sp.csr_matrix((z[p + "val"], z[p + "col"], z[p + "rowptr"]))
    M = H.CudaCSRMatrixInterface(ctx, z[p + "rowptr"], z[p + "col"], z[p + "val"])
    x = z[p + "x"]
    X = vec(ctx, x)
    Y = H.CudaVectorInterface(ctx, len(x))
    M.vmult_transpose(Y, X, 1.0, 0.0)
    assert np.allclose(Y.copy_back(), A.T @ x, rtol=1e-13, atol=1e-14)


def test_host_only_matrices(ctx):
    """gh_csr_create_host: validated like gh_csr_create, refused by every
    device operation (no device layout exists), accepted by the ordered
    solver, whose solve matches the solver built from device matrices"""
    import ctypes as C
    from gascoignelib_amd import native as N
    from gascoignelib_amd.native import GhError
    with pytest.raises(GhError):
        H.CudaCSRMatrixInterface.host_only(ctx, [0, 1, 2], [0, 5], [1.0, 1.0])
    rowptr, col, val = [0, 2, 4], [0, 1, 0, 1], [2.0, -1.0, -1.0, 2.0]
    A = H.CudaCSRMatrixInterface.host_only(ctx, rowptr, col, val)
    assert A.n() == 2 and A.n_entries() == 4
    X, Y, B = vec(ctx, np.ones(2)), H.CudaVectorInterface(ctx, 2), vec(ctx, np.ones(2))
    for call in (lambda: A.vmult(Y, X), lambda: A.vmult_transpose(Y, X),
                 lambda: A.residual(Y, X, B), lambda: A.smooth(Y, B, X, 1, 0.8),
                 lambda: A.jacobi(Y), lambda: A.layout()):
        with pytest.raises(GhError):
            call()
    nodes = np.array([[0, 1]], np.int32)
    V = C.c_void_p()
    assert N.hip().gh_vanka_create(ctx.h, A.h, 1, 2, N.ptr(nodes, N.i32p), 0, C.byref(V)) != 0
    p = N.GhMgParams()
    N.hip().gh_mg_params_default(C.byref(p))
    arr = (C.c_void_p * 1)(A.h)
    none = (C.c_void_p * 1)(None)
    mg = C.c_void_p()
    assert N.hip().gh_mg_create(ctx.h, 1, arr, none, none, none, C.byref(p), C.byref(mg)) != 0
    # ordered solvers run on host-only levels (the Python wrapper's choice)
    h = Hierarchy.from_fixture(fixture_path("q1_3d_16"))
    f = rhs(h)
    out = []
    for order in (None, [np.arange(lev.n) for lev in h.levels]):
        mgs = H.CudaMultiLevelSolver(ctx, h, ndirect=1, order=order, **MGP)
        U = mgs.new_vector()
        mgs.Solve(U, vec(ctx, f), H.CGInfo(1e-10, 1e-30, 100), H.NLInfo(1e-10, 1e-30, 30))
        out.append(U.copy_back())
        mgs.free()
    assert np.array_equal(out[0], out[1])


def test_invalid_arguments_are_errors(ctx):
    from gascoignelib_amd.native import GhError
    with pytest.raises(GhError):
        H.CudaCSRMatrixInterface(ctx, [0, 1, 2], [0, 5], [1.0, 1.0])  # column out of range
    with pytest.raises(GhError):
        H.CudaCSRMatrixInterface(ctx, [0, 1, 2], [1, 0], [1.0, 1.0])  # no diagonal
    with pytest.raises(GhError):
        H.HNStructure(ctx, 13, 4, [[3, 0, 9, 0]], np.zeros((0, 10)))  # parent out of range
    with pytest.raises(GhError):
        H.DirichletMask(ctx, 4, 1, [7], [0])
    A = H.CudaCSRMatrixInterface(ctx, [0, 1, 2], [0, 1], [2.0, 2.0])
    X, Y = H.CudaVectorInterface(ctx, 2), H.CudaVectorInterface(ctx, 3)
    with pytest.raises(GhError):
        A.vmult(Y, X)


# --------------------------------------------------------------- solver parity
def rhs(h, seed=7):
    lev = h.fine
    f = splitmix64_uniform(seed, lev.n * lev.ncomp)
    O.HN(lev.hn_kind, lev.hn_edges, lev.hn_faces).distribute(f)
    f[lev.dirichlet] = 0.0
    return f


MGP = dict(mgtype="V", iterpre=2, iterpost=2, iterexact=10, omega=0.8, mgomega=1.0)


def relerr(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("name", FIXTURES)
@pytest.mark.parametrize("ndirect", [1, 200])
@pytest.mark.parametrize("mgtype", ["V", "W", "F"])
def test_precondition_vs_oracle(ctx, name, ndirect, mgtype):
    print(f"[{name} ndirect={ndirect} {mgtype}]", flush=True)
    h = Hierarchy.from_fixture(fixture_path(name))
    params = dict(MGP, mgtype=mgtype)
    mg = H.CudaMultiLevelSolver(ctx, h, ndirect=ndirect, **params)
    om = O.MG(h, ndirect=ndirect, **params)
    f = rhs(h)
    xo = np.zeros_like(f)
    om.precondition(xo, f)
    X = mg.new_vector()
    mg.precondition(X, vec(ctx, f), x_is_zero=True)
    if ndirect == 1:
        assert np.array_equal(X.copy_back(), xo)
    assert relerr(X.copy_back(), xo) <= 1e-12
    assert mg.mgsteps == om.mgsteps
    # from a non-zero initial guess
    x0 = splitmix64_uniform(3, len(f))
    xo = x0.copy()
    om.precondition(xo, f)
    X.assign(x0)
    mg.precondition(X, vec(ctx, f), x_is_zero=False)
    assert relerr(X.copy_back(), xo) <= 1e-12


@pytest.mark.parametrize("name", ["q1_2d_128", "q1_3d_8_ada"])
def test_linear_mg_iterations_vs_oracle(ctx, name):
    h = Hierarchy.from_fixture(fixture_path(name))
    mg = H.CudaMultiLevelSolver(ctx, h, ndirect=1, **MGP)
    om = O.MG(h, ndirect=1, **MGP)
    f = rhs(h)
    io = O.cginfo(1e-6, 1e-30, 50)
    uo = np.zeros_like(f)
    om.linear_mg(uo, f, io)
    ig = H.CGInfo(1e-6, 1e-30, 50)
    U = mg.new_vector()
    mg.LinearMg(U, vec(ctx, f), ig)
    # the V-cycles are bit-exact and the stopping test is a ratio of norms
    # (sums in another order, a few ulps apart): the counts are equal unless a
    # ratio falls within rounding of the tolerance, which these cases do not
    assert ig.iteration == io.iteration
    assert ig.status == O.CG_STATUS[io.status]
    assert relerr(U.copy_back(), uo) <= 1e-12


@pytest.mark.parametrize("name", FIXTURES)
def test_gmres_vs_oracle(ctx, name):
    h = Hierarchy.from_fixture(fixture_path(name))
    mg = H.CudaMultiLevelSolver(ctx, h, ndirect=1, **MGP)
    om = O.MG(h, ndirect=1, **MGP)
    f = rhs(h)
    io = O.cginfo(1e-8, 1e-30, 100)
    xo = np.zeros_like(f)
    om.gmres(xo, f, io, vmax=10)
    ig = H.CGInfo(1e-8, 1e-30, 100)
    X = mg.new_vector()
    mg.Gmres(X, vec(ctx, f), ig, vmax=10)
    assert ig.iteration == io.iteration
    assert relerr(X.copy_back(), xo) <= 1e-12
    assert ig.residual == pytest.approx(io.residual, rel=1e-9)


@pytest.mark.parametrize("name", FIXTURES)
def test_newton_solve_vs_oracle(ctx, name):
    """StdMultiLevelSolver::Solve (Newton + GMRES(10) + V(2,2) Jacobi) to 1e-10."""
    print(f"[{name}]", flush=True)
    h = Hierarchy.from_fixture(fixture_path(name))
    mg = H.CudaMultiLevelSolver(ctx, h, ndirect=1, **MGP)
    om = O.MG(h, ndirect=1, **MGP)
    f = rhs(h)
    uo = np.zeros_like(f)
    rc, rep_o = om.newton(uo, f, O.cginfo(1e-10, 1e-30, 100), O.nlinfo(1e-10, 1e-30, 30))
    assert rc == 0
    U = mg.new_vector()
    rep = mg.Solve(U, vec(ctx, f), H.CGInfo(1e-10, 1e-30, 100), H.NLInfo(1e-10, 1e-30, 30))
    assert H.N.NL_STATUS[rep.status] == "converged"
    assert abs(rep.newton_iterations - rep_o.newton_iterations) <= 1
    assert abs(rep.gmres_iterations - rep_o.gmres_iterations) <= 1
    assert relerr(U.copy_back(), uo) <= 1e-12
    A = sp.csr_matrix((h.fine.val, h.fine.col, h.fine.rowptr))
    r = f - A @ U.copy_back()
    r[h.fine.dirichlet] = 0.0
    assert np.abs(r).max() <= 1e-10 * rep.first_residual * 1.0001


@pytest.mark.parametrize("prec", ["f32", "f16"])
@pytest.mark.parametrize("name", ["q1_2d_128", "q1_3d_16", "q1_3d_8_ada"])
def test_mixed_precision_solve(ctx, name, prec):
    """FP32 V-cycle (FP32 or FP16-stored values) inside FP64 GMRES + iterative
    refinement: same 1e-10 residual target, error vs the FP64 oracle stated."""
    h = Hierarchy.from_fixture(fixture_path(name))
    om = O.MG(h, ndirect=1, **MGP)
    f = rhs(h)
    uo = np.zeros_like(f)
    om.newton(uo, f, O.cginfo(1e-10, 1e-30, 100), O.nlinfo(1e-10, 1e-30, 30), norm_kind=1)
    mg = H.CudaMultiLevelSolver(ctx, h, ndirect=1, prec=prec, **MGP)
    U = mg.new_vector()
    rep = mg.Solve(U, vec(ctx, f), H.CGInfo(1e-10, 1e-30, 100), H.NLInfo(1e-10, 1e-30, 30),
                   norm_kind=1)
    assert H.N.NL_STATUS[rep.status] == "converged"
    u = U.copy_back()
    A = sp.csr_matrix((h.fine.val, h.fine.col, h.fine.rowptr))
    r = f - A @ u
    r[h.fine.dirichlet] = 0.0
    assert np.linalg.norm(r) <= 1e-10 * rep.first_residual * 1.0001
    err = relerr(u, uo)
    print(f"{name} {prec}: newton {rep.newton_iterations} gmres {rep.gmres_iterations} "
          f"rel. error vs FP64 oracle {err:.2e}")
    assert err <= 1e-8


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_graph_replay_matches_eager(ctx, prec):
    h = Hierarchy.from_fixture(fixture_path("q1_3d_16"))
    eager = H.CudaMultiLevelSolver(ctx, h, ndirect=100, prec=prec, **MGP)
    graph = H.CudaMultiLevelSolver(ctx, h, ndirect=100, prec=prec, use_graph=True, **MGP)
    X1, X2 = eager.new_vector(), graph.new_vector()
    for seed in range(4):
        f = rhs(h, seed)
        F = vec(ctx, f)
        eager.precondition(X1, F, x_is_zero=True)
        graph.precondition(X2, F, x_is_zero=True)
        # eager device work and a host round trip between replays
        X1.add(0.0, X2)
        assert np.array_equal(X1.copy_back(), X2.copy_back())
    assert graph.mgsteps == eager.mgsteps


@pytest.mark.parametrize("name", FIXTURES)
def test_ordered_solver_matches_unordered(ctx, name):
    """gh_mg_create_ordered renumbers every level for locality but keeps each
    row's summation order: with the iterative coarse solve the V-cycle is the
    same bit for bit, and whole solves agree with the oracle to 1e-12."""
    h = Hierarchy.from_fixture(fixture_path(name))
    f = rhs(h)
    om = O.MG(h, ndirect=1, **MGP)
    xo = np.zeros_like(f)
    om.precondition(xo, f)
    mg = H.CudaMultiLevelSolver(ctx, h, ndirect=1, order="lexicographic", **MGP)
    X = mg.new_vector()
    mg.precondition(X, vec(ctx, f), x_is_zero=True)
    assert np.array_equal(X.copy_back(), xo)
    uo = np.zeros_like(f)
    om.newton(uo, f, O.cginfo(1e-10, 1e-30, 100), O.nlinfo(1e-10, 1e-30, 30))
    U = mg.new_vector()
    rep = mg.Solve(U, vec(ctx, f), H.CGInfo(1e-10, 1e-30, 100), H.NLInfo(1e-10, 1e-30, 30))
    assert H.N.NL_STATUS[rep.status] == "converged"
    assert relerr(U.copy_back(), uo) <= 1e-12


@pytest.mark.parametrize("prec", ["f64", "f16"])
def test_ordered_graph_mixed(ctx, prec):
    h = Hierarchy.from_fixture(fixture_path("q1_3d_16"))
    a = H.CudaMultiLevelSolver(ctx, h, ndirect=100, prec=prec, **MGP)
    b = H.CudaMultiLevelSolver(ctx, h, ndirect=100, prec=prec, order="lexicographic",
                               use_graph=True, **MGP)
    f = rhs(h)
    F = vec(ctx, f)
    Ua, Ub = a.new_vector(), b.new_vector()
    ra = a.Solve(Ua, F, H.CGInfo(1e-6, 1e-30, 100), H.NLInfo(1e-10, 1e-30, 30), norm_kind=1)
    rb = b.Solve(Ub, F, H.CGInfo(1e-6, 1e-30, 100), H.NLInfo(1e-10, 1e-30, 30), norm_kind=1)
    assert H.N.NL_STATUS[rb.status] == "converged"
    assert abs(ra.gmres_iterations - rb.gmres_iterations) <= 1
    assert relerr(Ub.copy_back(), Ua.copy_back()) <= 1e-9


def sell_test_matrix(n, seed, equal_lengths):
    """rows of slices 2 mod 3 couple to far columns (those slices need 32-bit
    indices); the others are banded: slices 0 mod 3 on a 7-point band (shared
    patterns), slices 1 mod 3 on that band plus one random column within 2000
    (16-bit offsets from the slice base).  The pattern holds exact zeros,
    first entries included.  equal_lengths: every row keeps 5 entries (the
    others of its off-diagonal entries are zero) -- else a random fifth of
    all entries is zero"""
    rng = np.random.default_rng(seed)
    band = np.array([-300, -17, -1, 0, 1, 17, 300])
    rows = []
    for i in range(n):
        kind = (i // 64) % 3
        if kind == 2:
            cols = np.unique(np.concatenate([[i], rng.integers(0, n, 6)]))
        elif kind == 1:
            cols = np.unique(np.clip(np.append(i + band, i + rng.integers(-2000, 2000)), 0, n - 1))
        else:
            cols = np.unique(np.clip(i + band, 0, n - 1))
        rows.append(cols)
    rowptr = np.zeros(n + 1, np.int32)
    rowptr[1:] = np.cumsum([len(c) for c in rows])
    col = np.concatenate(rows).astype(np.int32)
    val = rng.uniform(-1, 1, len(col))
    for i, c in enumerate(rows):
        off = np.flatnonzero(c != i)
        if equal_lengths:
            drop = rng.choice(off, max(0, len(c) - 5), replace=False)
            if i % 5 == 0 and off[0] == 0 and len(drop) and 0 not in drop:
                drop[0] = 0  # a first entry
            val[rowptr[i] + drop] = 0.0
        else:
            val[rowptr[i] + off[rng.random(len(off)) < 0.2]] = 0.0
            if i % 5 == 0 and off[0] == 0:
                val[rowptr[i]] = 0.0
        val[rowptr[i] + np.flatnonzero(c == i)] = 4.0
    return rowptr, col, val


def check_products_vs_oracle(ctx, A, rowptr, col, val, seed):
    rng = np.random.default_rng(seed)
    n = len(rowptr) - 1
    Ao = O.Matrix(rowptr, col, val)
    x, b = rng.uniform(-1, 1, n), rng.uniform(-1, 1, n)
    y = np.zeros(n)
    Ao.vmult(y, x)
    Y = H.CudaVectorInterface(ctx, n)
    A.vmult(Y, vec(ctx, x), 1.0, 0.0)
    assert np.array_equal(Y.copy_back(), y)
    r = np.zeros(n)
    Ao.residual(r, x, b)
    R = H.CudaVectorInterface(ctx, n)
    A.residual(R, vec(ctx, x), vec(ctx, b))
    assert np.array_equal(R.copy_back(), r)


def test_sell_column_encodings_and_dropped_zeros(ctx):
    """A matrix whose slices mix 16-bit, 32-bit and pattern column encodings
    and whose pattern holds exact zeros, with rows of one kept length (so the
    rows run in their own order): SpMV and residual equal the oracle's (same
    values, same per-row order)."""
    rowptr, col, val = sell_test_matrix(150_000, 5, equal_lengths=True)
    A = H.CudaCSRMatrixInterface(ctx, rowptr, col, val)
    lay = A.layout()
    assert lay["stored"] == int(np.count_nonzero(val))
    assert not lay["permuted"]
    assert 0 < lay["narrow"] < lay["slices"]
    assert 0 < lay["patterned"] < lay["slices"] - lay["narrow"]  # and some 32-bit slices
    check_products_vs_oracle(ctx, A, rowptr, col, val, 6)


def test_sell_rows_grouped_by_length(ctx):
    """Rows of varying kept length (a random fifth of the entries are exact
    zeros) run in slices of equal-length rows: fewer padded slots than the
    rows' own order needs, the same bits as the oracle's row sums"""
    rowptr, col, val = sell_test_matrix(150_000, 7, equal_lengths=False)
    A = H.CudaCSRMatrixInterface(ctx, rowptr, col, val)
    lay = A.layout()
    kept = np.add.reduceat(val != 0, rowptr[:-1])
    pad = np.zeros(lay["slices"] * 64, np.int64)
    pad[: len(kept)] = kept
    own = 64 * int(pad.reshape(-1, 64).max(1).sum())
    assert lay["permuted"] and lay["stored"] == int(np.count_nonzero(val))
    assert lay["slots"] * 20 <= own * 19
    check_products_vs_oracle(ctx, A, rowptr, col, val, 8)


def long_row_matrix(n, L, seed):
    """Rows of up to L entries in four kinds of slices: one band of L offsets
    for the whole slice, two bands alternating by row, random columns within
    2000 of the row (16-bit offsets), random columns anywhere (32-bit indices);
    the last two vary the row length and hold exact zeros (dropped from the
    layout).  Values are nonzero multiples of 2^-10 in [-1, 1] elsewhere and
    the diagonal is L + 1, all exact in f16.  The band rows keep every entry,
    so their slices run from the pattern table in the rows' own order."""
    rng = np.random.default_rng(seed)
    band_a = np.arange(-(L // 2), L - L // 2) * 3
    band_b = np.arange(-(L // 2), L - L // 2) * 5
    rows = []
    for i in range(n):
        kind = (i // 64) % 4
        if kind == 0:
            c = i + band_a
        elif kind == 1:
            c = i + (band_a if i % 2 == 0 else band_b)
        elif kind == 2:
            c = i + rng.integers(-2000, 2000, rng.integers(1, L))
        else:
            c = rng.integers(0, n, rng.integers(1, L))
        rows.append(np.unique(np.append(np.clip(c, 0, n - 1), i)))
    rowptr = np.zeros(n + 1, np.int32)
    rowptr[1:] = np.cumsum([len(c) for c in rows])
    col = np.concatenate(rows).astype(np.int32)
    val = rng.choice([-1.0, 1.0], len(col)) * rng.integers(1, 1025, len(col)) / 1024.0
    for i, c in enumerate(rows):
        if (i // 64) % 4 >= 2:
            val[rowptr[i] + np.flatnonzero(rng.random(len(c)) < 0.05)] = 0.0
        val[rowptr[i] + np.flatnonzero(c == i)] = L + 1.0
    return rowptr, col, val


@pytest.mark.parametrize("prec", ["f32", "f16"])
@pytest.mark.parametrize("n,L", [(5000, 12), (5000, 24), (5000, 48), (5000, 100), (5000, 200),
                                 (90, 30), (100, 48), (700, 200), (70000, 100)])
def test_reduced_precision_sweeps_long_rows(ctx, n, L, prec):
    """FP32-arithmetic products, residuals and Jacobi / Richardson sweeps on
    levels of few slices and long rows, where one slice's entries are split
    over up to 16 wavefronts and the partial sums meet in LDS: each result is
    within the FP32 summation bound (L + 20) 2^-23 sum|terms| of the exact
    value (the stored values are exact in f16 and f32), rows past the last
    full slice included."""
    print(f"[n={n} L={L} {prec}]", flush=True)
    rowptr, col, val = long_row_matrix(n, L, seed=L + n)
    M = sp.csr_matrix((val, col, rowptr), shape=(n, n))
    absM = abs(M)
    A = H.CudaCSRMatrixInterface(ctx, rowptr, col, val, prec=prec)
    lay = A.layout()
    assert lay["stored"] == int(np.count_nonzero(val))
    if n >= 5000:
        assert lay["patterned"] > 0 and lay["narrow"] > 0
    if n > 65536:  # columns anywhere need 32-bit indices
        assert lay["patterned"] + lay["narrow"] < lay["slices"]
    rng = np.random.default_rng(n * L)
    f32 = lambda v: v.astype(np.float32).astype(np.float64)  # noqa: E731
    x, b, y0 = (f32(rng.uniform(-1, 1, n)) for _ in range(3))
    vx = H.CudaVectorInterface.from_global(ctx, x, 1, "f32")
    vb = H.CudaVectorInterface.from_global(ctx, b, 1, "f32")
    eps = (L + 20) * 2.0 ** -23
    ax, sx = M @ x, absM @ np.abs(x)

    Y = H.CudaVectorInterface.from_global(ctx, y0, 1, "f32")
    A.vmult(Y, vx, -0.5, 1.5)
    want = -0.5 * ax + 1.5 * y0
    assert np.all(np.abs(Y.copy_back() - want) <= eps * (0.5 * sx + 1.5 * np.abs(y0)) + 1e-30)

    R = H.CudaVectorInterface(ctx, n, 1, "f32")
    A.residual(R, vx, vb)
    r = b - ax
    bound_r = eps * (sx + np.abs(b))
    assert np.all(np.abs(R.copy_back() - r) <= bound_r + 1e-30)

    dinv = 1.0 / M.diagonal()
    for smoother, scale in (("jacobi", 0.8 * dinv), ("richardson", np.full(n, 0.8))):
        X = H.CudaVectorInterface.from_global(ctx, x, 1, "f32")
        T = H.CudaVectorInterface(ctx, n, 1, "f32")
        A.smooth(X, vb, T, 1, 0.8, smoother)
        want = x + scale * r
        bound = scale * bound_r + 2.0 ** -21 * (np.abs(x) + np.abs(scale * r))
        assert np.all(np.abs(X.copy_back() - want) <= bound), smoother


@pytest.mark.parametrize("prec", ["f64", "f32", "f16"])
@pytest.mark.parametrize("name", ["q1_2d_128", "q1_3d_16", "q1_3d_8_ada", "q2_3d_8_ada"])
def test_flexible_gmres_solve(ctx, name, prec):
    """Flexible GMRES around the V-cycle (FP64 or lower precision): the true FP64
    residual reaches the 1e-10 target in one outer step, and the solution agrees
    with the FP64 oracle solve."""
    print(f"[{name} {prec}]", flush=True)
    h = Hierarchy.from_fixture(fixture_path(name))
    om = O.MG(h, ndirect=1, **MGP)
    f = rhs(h)
    uo = np.zeros_like(f)
    om.newton(uo, f, O.cginfo(1e-12, 1e-30, 100), O.nlinfo(1e-12, 1e-30, 30), norm_kind=1)
    mg = H.CudaMultiLevelSolver(ctx, h, ndirect=1, prec=prec, krylov="fgmres", **MGP)
    U = mg.new_vector()
    rep = mg.Solve(U, vec(ctx, f), H.CGInfo(0.5e-10, 1e-30, 100), H.NLInfo(1e-10, 1e-30, 30),
                   norm_kind=1)
    assert H.N.NL_STATUS[rep.status] == "converged"
    u = U.copy_back()
    A = sp.csr_matrix((h.fine.val, h.fine.col, h.fine.rowptr))
    r = f - A @ u
    r[h.fine.dirichlet] = 0.0
    assert np.linalg.norm(r) <= 1e-10 * rep.first_residual * 1.0001
    err = relerr(u, uo)
    print(f"{name} {prec}: newton {rep.newton_iterations} fgmres {rep.gmres_iterations} "
          f"rel. error vs FP64 oracle {err:.2e}")
    assert err <= 1e-8


@pytest.mark.parametrize("prec", ["f64", "f16"])
def test_fgmres_tolerance_base_and_estimate(ctx, prec):
    """Flexible GMRES measures its relative tolerance from the residual it
    starts from, and its residual estimate is the true FP64 residual (the
    products A z_j are FP64); the reference's left-preconditioned GMRES instead
    takes the residual after its first step as the base"""
    h = Hierarchy.from_fixture(fixture_path("q1_3d_16"))
    lev = h.fine
    f = rhs(h)
    A = sp.csr_matrix((lev.val, lev.col, lev.rowptr), shape=(lev.n, lev.n))
    r0 = f.copy()
    r0[lev.dirichlet] = 0.0
    mg = H.CudaMultiLevelSolver(ctx, h, ndirect=1, prec=prec, krylov="fgmres", **MGP)
    info = H.CGInfo(1e-6, 1e-30, 50)
    X = mg.new_vector()
    assert mg.Gmres(X, vec(ctx, f), info, vmax=20)
    r = f - A @ X.copy_back()
    r[lev.dirichlet] = 0.0
    assert info.s.firstresidual == pytest.approx(np.linalg.norm(r0), rel=1e-12)
    assert info.s.residual == pytest.approx(np.linalg.norm(r), rel=1e-6)
    assert np.linalg.norm(r) <= 1e-6 * np.linalg.norm(r0)
    # the left-preconditioned GMRES works on M^-1 r; its first recorded
    # residual is the one after its first step, well below |M^-1 r0|
    mg64 = H.CudaMultiLevelSolver(ctx, h, ndirect=1, **MGP)
    info64 = H.CGInfo(1e-6, 1e-30, 50)
    mg64.Gmres(mg64.new_vector(), vec(ctx, f), info64, vmax=10)
    Z = mg64.new_vector()
    mg64.precondition(Z, vec(ctx, r0), x_is_zero=True)
    assert info64.s.firstresidual < 0.5 * np.linalg.norm(Z.copy_back())
