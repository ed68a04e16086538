"""Q2 patch-element Poisson (CGDiscQ23d, up to 125 couplings per row) on the
GPU against the oracle: the V-cycle bit for bit (every level smoothed) and
within 1e-12 with a direct coarse level, Newton/GMRES iterates to 1e-12 with
iteration counts within one, the mixed-precision solve within 1e-8 of the
FP64 oracle solve, and the adjoint solve against the oracle on the transposed
levels."""
import numpy as np
import pytest

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
