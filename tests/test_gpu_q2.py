"""Q2 patch-element Poisson (CGDiscQ23d, up to 125 couplings per row) on the
GPU against the oracle: the V-cycle bit for bit (every level smoothed) and
within 1e-12 with a direct coarse level, Newton/GMRES iterates to 1e-12 with
iteration counts within one, the mixed-precision solve within 1e-8 of the
FP64 oracle solve, and the adjoint solve against the oracle on the transposed
levels."""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle_ctypes as O
from gascoignelib_amd import hip as H
from gascoignelib_amd.problem import lattice_rhs
from gascoignelib_amd.structure import q2_hierarchy

pytestmark = pytest.mark.gpu

MGP = dict(mgtype="V", iterpre=2, iterpost=2, iterexact=10, omega=0.8, mgomega=1.0)


@pytest.fixture(scope="module")
def ctx():
    c = H.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def h8():
    return q2_hierarchy(3)


@pytest.fixture(scope="module")
def h16():
    return q2_hierarchy(4)


def rhs(h, seed=3):
    f = lattice_rhs(h.fine.coords, seed)
    f[h.fine.dirichlet] = 0.0
    return f


@pytest.mark.parametrize("ndirect", [1, 200])
@pytest.mark.parametrize("order", [None, "lexicographic"])
def test_q2_vcycle_vs_oracle(ctx, h8, order, ndirect):
    print(f"[q2 V-cycle 8^3 order={order} ndirect={ndirect}]", flush=True)
    f = rhs(h8)
    om = O.MG(h8, ndirect=ndirect, **MGP)
    xo = np.zeros_like(f)
    om.precondition(xo, f)
    mg = H.CudaMultiLevelSolver(ctx, h8, ndirect=ndirect, order=order, **MGP)
    X = mg.new_vector()
    mg.precondition(X, H.CudaVectorInterface.from_global(ctx, f))
    x = X.copy_back()
    if ndirect == 1:
        assert np.array_equal(x, xo)
    assert np.linalg.norm(x - xo) <= 1e-12 * np.linalg.norm(xo)


def test_q2_solve_vs_oracle(ctx, h16):
    print("[q2 Newton/GMRES 16^3 FP64 and mixed]", flush=True)
    f = rhs(h16)
    om = O.MG(h16, ndirect=200, **MGP)
    uo = np.zeros_like(f)
    rc, orep = om.newton(uo, f, O.cginfo(1e-10, 1e-30, 100), O.nlinfo(1e-10, 1e-30, 30),
                         norm_kind=1)
    assert rc == 0
    mg = H.CudaMultiLevelSolver(ctx, h16, ndirect=200, order="lexicographic", **MGP)
    U = mg.new_vector()
    F = H.CudaVectorInterface.from_global(ctx, f)
    rep = mg.Solve(U, F, H.CGInfo(1e-10, 1e-30, 100), H.NLInfo(1e-10, 1e-30, 30), norm_kind=1)
    assert rep.newton_iterations == orep.newton_iterations
    assert abs(rep.gmres_iterations - orep.gmres_iterations) <= 1
    assert np.linalg.norm(U.copy_back() - uo) <= 1e-12 * np.linalg.norm(uo)
    # mixed precision: f16 values, FP32 V-cycle inside FP64 flexible GMRES
    uo2 = np.zeros_like(f)
    om.newton(uo2, f, O.cginfo(1e-12, 1e-30, 100), O.nlinfo(1e-12, 1e-30, 30), norm_kind=1)
    mx = H.CudaMultiLevelSolver(ctx, h16, ndirect=200, order="lexicographic", prec="f16",
                                krylov="fgmres", **MGP)
    U2 = mx.new_vector()
    rep2 = mx.Solve(U2, F, H.CGInfo(0.9e-10, 1e-30, 100), H.NLInfo(1e-10, 1e-30, 30),
                    norm_kind=1)
    assert H.N.NL_STATUS[rep2.status] == "converged"
    assert np.linalg.norm(U2.copy_back() - uo2) <= 1e-8 * np.linalg.norm(uo2)


def test_q2_adjoint_vs_oracle(ctx, h16):
    """A is symmetric here, so the adjoint solve must give the forward answer
    for the same right-hand side; checked against the oracle's forward solve"""
    print("[q2 adjoint 16^3]", flush=True)
    g = rhs(h16, seed=9)
    om = O.MG(h16, ndirect=200, **MGP)
    zo = np.zeros_like(g)
    om.newton(zo, g, O.cginfo(1e-12, 1e-30, 100), O.nlinfo(1e-12, 1e-30, 30), norm_kind=1)
    mg = H.CudaMultiLevelSolver(ctx, h16, ndirect=200, order="lexicographic", adjoint=True, **MGP)
    Z = mg.new_vector()
    mg.Solve(Z, H.CudaVectorInterface.from_global(ctx, g), H.CGInfo(1e-12, 1e-30, 100),
             H.NLInfo(1e-12, 1e-30, 30), norm_kind=1)
    assert np.linalg.norm(Z.copy_back() - zo) <= 1e-10 * np.linalg.norm(zo)


def test_level_matrix_view(ctx, h8):
    """gh_mg_level_matrix: the ordered solver's FP64 level matrix is the
    caller's matrix renumbered with each row's entries in the caller's order
    (same sums, same bits); the f16 V-cycle's matrix agrees to f16 rounding"""
    from gascoignelib_amd.problem import lexicographic_order
    print("[level matrix view]", flush=True)
    lev = h8.fine
    lex = lexicographic_order(lev.coords)
    x = np.random.default_rng(2).uniform(-1, 1, lev.n)
    A = H.CudaCSRMatrixInterface(ctx, lev.rowptr, lev.col, lev.val)
    Y = H.CudaVectorInterface(ctx, lev.n)
    A.vmult(Y, H.CudaVectorInterface.from_global(ctx, x))
    y = Y.copy_back()
    for prec, fp64, tol in (("f64", True, 0.0), ("f16", False, 2e-3)):
        mg = H.CudaMultiLevelSolver(ctx, h8, ndirect=200, order="lexicographic", prec=prec,
                                    krylov="fgmres" if prec != "f64" else "gmres", **MGP)
        Ai = mg.level_matrix(fp64=fp64)
        assert Ai.n() == lev.n
        vp = "f64" if fp64 else "f32"
        Yi = H.CudaVectorInterface(ctx, lev.n, 1, vp)
        Ai.vmult(Yi, H.CudaVectorInterface.from_global(ctx, x[lex], 1, vp))
        yi = Yi.copy_back()
        if tol == 0.0:
            assert np.array_equal(yi, y[lex])
        else:
            assert np.abs(yi - y[lex]).max() <= tol * np.abs(y).max()
        del Ai
        mg.free()


# ------------------------------------------------- adaptive levels, Q2 hanging nodes
ADA = dict(adaptive=2, radius=0.25, centre=(0.5, 0.47, 0.53))


@pytest.fixture(scope="module")
def a8():
    return q2_hierarchy(3, **ADA)


@pytest.fixture(scope="module")
def a16():
    return q2_hierarchy(4, **ADA)


def ada_rhs(h, seed=3):
    """seeded f distributed onto the parents of the hanging nodes
    (HNDistribute, as the reference's assembly leaves it), Dirichlet 0"""
    lev = h.fine
    f = lattice_rhs(lev.coords, seed)
    O.HN(lev.hn_kind, lev.hn_edges, lev.hn_faces).distribute(f)
    f[lev.dirichlet] = 0.0
    return f


@pytest.mark.parametrize("ndirect", [1, 200])
@pytest.mark.parametrize("order", [None, "lexicographic"])
def test_q2_adaptive_vcycle_vs_oracle(ctx, a8, order, ndirect):
    """8^3 + 2 adaptive levels: the V-cycle with HNAverage / HNDistribute /
    HNZero on both adaptive levels, bit for bit with the iterative coarse
    level, from a zero and from a non-zero start"""
    print(f"[q2 adaptive V-cycle 8^3+2 order={order} ndirect={ndirect}]", flush=True)
    f = ada_rhs(a8)
    om = O.MG(a8, ndirect=ndirect, **MGP)
    mg = H.CudaMultiLevelSolver(ctx, a8, ndirect=ndirect, order=order, **MGP)
    x0 = np.random.default_rng(4).uniform(-1, 1, len(f))
    for start in (np.zeros_like(f), x0):
        xo = start.copy()
        om.precondition(xo, f)
        X = H.CudaVectorInterface.from_global(ctx, start)
        mg.precondition(X, H.CudaVectorInterface.from_global(ctx, f),
                        x_is_zero=not start.any())
        x = X.copy_back()
        if ndirect == 1:
            assert np.array_equal(x, xo)
        assert np.linalg.norm(x - xo) <= 1e-12 * np.linalg.norm(xo)
    assert mg.mgsteps == om.mgsteps


def test_q2_adaptive_solve_vs_oracle(ctx, a16):
    """16^3 + 2 adaptive levels (77,485 DOF, 6740 edge and 3352 face hanging
    nodes): Newton/GMRES iterates to 1e-12 with iteration counts within one,
    the true residual of the condensed system, and the mixed-precision solve
    within 1e-8 of the FP64 oracle solve"""
    print("[q2 adaptive Newton/GMRES 16^3+2 FP64 and mixed]", flush=True)
    f = ada_rhs(a16)
    lev = a16.fine
    om = O.MG(a16, ndirect=200, **MGP)
    uo = np.zeros_like(f)
    rc, orep = om.newton(uo, f, O.cginfo(1e-10, 1e-30, 100), O.nlinfo(1e-10, 1e-30, 30),
                         norm_kind=1)
    assert rc == 0
    mg = H.CudaMultiLevelSolver(ctx, a16, ndirect=200, order="lexicographic", **MGP)
    U = mg.new_vector()
    F = H.CudaVectorInterface.from_global(ctx, f)
    rep = mg.Solve(U, F, H.CGInfo(1e-10, 1e-30, 100), H.NLInfo(1e-10, 1e-30, 30), norm_kind=1)
    assert rep.newton_iterations == orep.newton_iterations
    assert abs(rep.gmres_iterations - orep.gmres_iterations) <= 1
    u = U.copy_back()
    assert np.linalg.norm(u - uo) <= 1e-12 * np.linalg.norm(uo)
    A = sp.csr_matrix((lev.val, lev.col, lev.rowptr), shape=(lev.n, lev.n))
    r = f - A @ u
    r[lev.dirichlet] = 0.0
    assert np.linalg.norm(r) <= 1e-10 * rep.first_residual * 1.0001
    # the hanging entries leave the solve zeroed (HNZero)
    assert not u[lev.hn_edges[:, 0]].any() and not u[lev.hn_faces[:, 0]].any()
    uo2 = np.zeros_like(f)
    om.newton(uo2, f, O.cginfo(1e-12, 1e-30, 100), O.nlinfo(1e-12, 1e-30, 30), norm_kind=1)
    mx = H.CudaMultiLevelSolver(ctx, a16, ndirect=200, order="lexicographic", prec="f16",
                                krylov="fgmres", **MGP)
    U2 = mx.new_vector()
    rep2 = mx.Solve(U2, F, H.CGInfo(0.9e-10, 1e-30, 100), H.NLInfo(1e-10, 1e-30, 30),
                    norm_kind=1)
    assert H.N.NL_STATUS[rep2.status] == "converged"
    assert np.linalg.norm(U2.copy_back() - uo2) <= 1e-8 * np.linalg.norm(uo2)


def test_q2_adaptive_adjoint_vs_oracle(ctx, a16):
    """the condensed operator is symmetric: the adjoint solve (transposed
    levels) of g equals the oracle's forward solve of g"""
    print("[q2 adaptive adjoint 16^3+2]", flush=True)
    g = ada_rhs(a16, seed=9)
    om = O.MG(a16, ndirect=200, **MGP)
    zo = np.zeros_like(g)
    om.newton(zo, g, O.cginfo(1e-12, 1e-30, 100), O.nlinfo(1e-12, 1e-30, 30), norm_kind=1)
    mg = H.CudaMultiLevelSolver(ctx, a16, ndirect=200, order="lexicographic", adjoint=True, **MGP)
    Z = mg.new_vector()
    mg.Solve(Z, H.CudaVectorInterface.from_global(ctx, g), H.CGInfo(1e-12, 1e-30, 100),
             H.NLInfo(1e-12, 1e-30, 30), norm_kind=1)
    assert np.linalg.norm(Z.copy_back() - zo) <= 1e-10 * np.linalg.norm(zo)


def column_scaled(h):
    """every level matrix scaled by columns, A_l D_l, with d = 1 + 0.4 sin(7x)
    cos(5y) sin(3z + 1) on free nodes and 1 on hanging and Dirichlet nodes:
    hanging rows and columns stay identity-only, and A_l D_l is not symmetric"""
    import dataclasses
    from gascoignelib_amd.problem import Hierarchy
    levels, ds = [], []
    for lev in h.levels:
        X = lev.coords
        d = 1.0 + 0.4 * np.sin(7 * X[:, 0]) * np.cos(5 * X[:, 1]) * np.sin(3 * X[:, 2] + 1)
        d[lev.dirichlet] = 1.0
        d[lev.hn_edges[:, 0]] = 1.0
        d[lev.hn_faces[:, 0]] = 1.0
        levels.append(dataclasses.replace(lev, val=lev.val * d[lev.col]))
        ds.append(d)
    return Hierarchy(levels, h.dim, h.name + ":scaled"), ds[-1]


def test_q2_adaptive_adjoint_nonsymmetric(ctx, a16):
    """adjoint with hanging nodes on a non-symmetric condensed operator:
    (A D)^T z = g on the free nodes is A z = D^-1 g there (A symmetric on the
    free block), so the GPU adjoint solve of the column-scaled hierarchy must
    equal the oracle's forward solve of the unscaled one with D^-1 g; the
    forward solve of A D against the oracle on the same scaled levels"""
    print("[q2 adaptive adjoint 16^3+2, column-scaled]", flush=True)
    hs, d = column_scaled(a16)
    lev = hs.fine
    A = sp.csr_matrix((lev.val, lev.col, lev.rowptr), shape=(lev.n, lev.n))
    assert abs(A - A.T).max() > 1e-3 * abs(A).max()
    g = ada_rhs(a16, seed=9)
    om = O.MG(a16, ndirect=200, **MGP)
    zo = np.zeros_like(g)
    om.newton(zo, g / d, O.cginfo(1e-12, 1e-30, 100), O.nlinfo(1e-12, 1e-30, 30), norm_kind=1)
    for order in (None, "lexicographic"):
        mg = H.CudaMultiLevelSolver(ctx, hs, ndirect=200, order=order, adjoint=True, **MGP)
        Z = mg.new_vector()
        rep = mg.Solve(Z, H.CudaVectorInterface.from_global(ctx, g), H.CGInfo(1e-12, 1e-30, 100),
                       H.NLInfo(1e-12, 1e-30, 30), norm_kind=1)
        assert H.N.NL_STATUS[rep.status] == "converged"
        assert np.linalg.norm(Z.copy_back() - zo) <= 1e-10 * np.linalg.norm(zo)
        mg.free()
    # forward on the scaled levels: bit for bit V-cycle, solve to 1e-12
    f = ada_rhs(a16)
    oms = O.MG(hs, ndirect=1, **MGP)
    xo = np.zeros_like(f)
    oms.precondition(xo, f)
    mg = H.CudaMultiLevelSolver(ctx, hs, ndirect=1, **MGP)
    X = mg.new_vector()
    mg.precondition(X, H.CudaVectorInterface.from_global(ctx, f), x_is_zero=True)
    assert np.array_equal(X.copy_back(), xo)
