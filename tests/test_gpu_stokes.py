"""Stokes-LPS channel on the GPU against the oracle.

The problem: NavierStokesLps<3> at U = 0 on the unit cube (components p, u,
v, w; 4x4 blocks on the patch structure of CGDiscQ13dLps), parabolic inflow
of seeded amplitude at x = 0, no-slip walls, do-nothing outflow at x = 1,
multigrid with the Vanka smoother on the 2x2x2-cell patches (108 local
unknowns) and a direct solve on the 2x2x2-cell level.

FP64: the V-cycle agrees with the oracle bit for bit when every level is
smoothed (the block products are taken in the reference's order, the patch
factors with the oracle's pivots) and within 1e-12 with a direct coarse level;
Newton/GMRES iterates to 1e-12, iteration counts within one.  Mixed precision
(FP32 Vanka V-cycle inside FP64 flexible GMRES): error against the FP64 oracle
solve within 1e-8.  Mass conservation is checked on every solve: the discrete
divergence rows of the constant test function sum to the boundary flux, and
the LPS term vanishes on constants, so the outflow flux equals the inflow
flux of the lifted data up to the solver tolerance.
"""
import numpy as np
import pytest

import oracle_ctypes as O
from gascoignelib_amd import hip as H
from gascoignelib_amd.structure import stokes_hierarchy

pytestmark = pytest.mark.gpu

MGP = dict(mgtype="V", iterpre=2, iterpost=2, iterexact=4, omega=0.8, mgomega=1.0, ndirect=200)


@pytest.fixture(scope="module")
def ctx():
    c = H.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def h8():
    return stokes_hierarchy(3, seed=5)  # 8^3 cells, 2,916 unknowns


@pytest.fixture(scope="module")
def h16():
    return stokes_hierarchy(4, seed=7)  # 16^3 cells, 19,652 unknowns


def flux_error(h, x):
    u = x.reshape(-1, 4) + h.u0
    fin = h.mass_in @ h.u0[:, 1]
    return abs(h.mass_out @ u[:, 1] - fin) / fin


@pytest.mark.parametrize("ndirect", [1, 200])
@pytest.mark.parametrize("order", [None, "lexicographic"])
def test_stokes_vcycle_vs_oracle(ctx, h8, order, ndirect):
    """ndirect 1: the 2x2x2-cell level is one Vanka patch, smoothed like the
    others, and the V-cycle is bit-exact; ndirect 200: that level is solved by
    the dense direct solve (LU in the oracle, an explicit inverse on the
    device), within 1e-12"""
    print(f"[stokes V-cycle 8^3 order={order} ndirect={ndirect}]", flush=True)
    f = h8.rhs
    p = dict(MGP, ndirect=ndirect)
    om = O.MG(h8, smoother="vanka", patches=h8.patches, **p)
    xo = np.zeros_like(f)
    om.precondition(xo, f)
    mg = H.CudaMultiLevelSolver(ctx, h8, linear_smooth="vanka", patches=h8.patches, order=order,
                                **p)
    X = mg.new_vector()
    mg.precondition(X, H.CudaVectorInterface.from_global(ctx, f, 4))
    x = X.copy_back()
    if ndirect == 1:
        assert np.array_equal(x, xo)
    assert np.linalg.norm(x - xo) <= 1e-12 * np.linalg.norm(xo)
    assert mg.mgsteps == om.mgsteps


def test_stokes_solve_vs_oracle(ctx, h16):
    print("[stokes Newton/GMRES 16^3 FP64]", flush=True)
    f = h16.rhs
    om = O.MG(h16, smoother="vanka", patches=h16.patches, **MGP)
    uo = np.zeros_like(f)
    rc, orep = om.newton(uo, f, O.cginfo(1e-10, 1e-30, 100), O.nlinfo(1e-10, 1e-30, 30),
                         norm_kind=1)
    assert rc == 0
    mg = H.CudaMultiLevelSolver(ctx, h16, linear_smooth="vanka", patches=h16.patches,
                                order="lexicographic", **MGP)
    U = mg.new_vector()
    rep = mg.Solve(U, H.CudaVectorInterface.from_global(ctx, f, 4), H.CGInfo(1e-10, 1e-30, 100),
                   H.NLInfo(1e-10, 1e-30, 30), norm_kind=1)
    u = U.copy_back()
    assert H.N.NL_STATUS[rep.status] == "converged"
    assert rep.newton_iterations == orep.newton_iterations
    assert abs(rep.gmres_iterations - orep.gmres_iterations) <= 1
    assert np.linalg.norm(u - uo) <= 1e-12 * np.linalg.norm(uo)
    assert flux_error(h16, u) < 1e-8


def test_stokes_mixed_precision(ctx, h16):
    print("[stokes FP32 Vanka V-cycle in FP64 FGMRES 16^3]", flush=True)
    f = h16.rhs
    om = O.MG(h16, smoother="vanka", patches=h16.patches, **MGP)
    uo = np.zeros_like(f)
    om.newton(uo, f, O.cginfo(1e-12, 1e-30, 100), O.nlinfo(1e-12, 1e-30, 30), norm_kind=1)
    mg = H.CudaMultiLevelSolver(ctx, h16, linear_smooth="vanka", patches=h16.patches,
                                order="lexicographic", prec="f32", krylov="fgmres", **MGP)
    U = mg.new_vector()
    rep = mg.Solve(U, H.CudaVectorInterface.from_global(ctx, f, 4),
                   H.CGInfo(0.9e-10, 1e-30, 100), H.NLInfo(1e-10, 1e-30, 30), norm_kind=1)
    assert H.N.NL_STATUS[rep.status] == "converged"
    u = U.copy_back()
    assert np.linalg.norm(u - uo) <= 1e-8 * np.linalg.norm(uo)
    assert flux_error(h16, u) < 1e-8
    # the FP64 residual of the returned iterate is what the outer loop certified
    A = H.CudaCSRMatrixInterface(ctx, h16.fine.rowptr, h16.fine.col, h16.fine.val, 4)
    R = H.CudaVectorInterface.from_global(ctx, f, 4)
    A.vmult(R, U, -1.0, 1.0)
    r = R.copy_back()
    r.reshape(-1, 4)[h16.fine.dirichlet, 1:] = 0.0
    assert np.linalg.norm(r) <= 1e-9 * np.linalg.norm(f)


@pytest.mark.parametrize("B", [2, 3, 4])
def test_block_products_bitexact(ctx, B):
    """Block products y = d A x + b y and r = b - A x on the component-row
    layout against the oracle's SparseBlockMatrix::vmult order, with block
    rows that are all zero (left out of the layout) and x in FP64 and FP32."""
    from gascoignelib_amd.problem import block_operator
    from gascoignelib_amd.structure import uniform_hierarchy
    print(f"[block products B={B}]", flush=True)
    lev = uniform_hierarchy("cube.inp", 3, 3, drop_coarse=1).fine
    rng = np.random.default_rng(B)
    Cm = rng.uniform(-1, 1, (B, B))
    Cm[B - 1] = 0.0  # last component couples through the diagonal block only
    S = rng.uniform(-0.5, 0.5, (B, B)) + 4 * np.eye(B)
    val = block_operator(lev, Cm, S)
    A = H.CudaCSRMatrixInterface(ctx, lev.rowptr, lev.col, val, B)
    Ao = O.Matrix(lev.rowptr, lev.col, val, B)
    m = lev.n * B
    x, y0, b = (rng.uniform(-1, 1, m) for _ in range(3))
    X = H.CudaVectorInterface.from_global(ctx, x, B)
    Y = H.CudaVectorInterface.from_global(ctx, y0, B)
    A.vmult(Y, X, 0.7, 1.0)
    yo = y0.copy()
    Ao.vmult(yo, x, 0.7)
    assert np.array_equal(Y.copy_back(), yo)
    R = H.CudaVectorInterface(ctx, lev.n, B)
    A.residual(R, X, H.CudaVectorInterface.from_global(ctx, b, B))
    ro = np.zeros(m)
    Ao.residual(ro, x, b)
    assert np.array_equal(R.copy_back(), ro)
    lay = A.layout()
    V = val.reshape(-1, B, B)
    assert lay["stored"] == int(np.count_nonzero(np.any(V != 0, axis=2)))
    assert lay["slots"] >= lay["stored"] and lay["slots"] % 64 == 0
    # FP32 values and vectors
    A32 = H.CudaCSRMatrixInterface(ctx, lev.rowptr, lev.col, val, B, "f32")
    X32 = H.CudaVectorInterface.from_global(ctx, x, B, "f32")
    Y32 = H.CudaVectorInterface(ctx, lev.n, B, "f32")
    A32.vmult(Y32, X32)
    ye = np.zeros(m)
    Ao.vmult(ye, x)
    assert np.abs(Y32.copy_back() - ye).max() <= 1e-5 * np.abs(ye).max()


def test_block_transpose_product(ctx, h8):
    """y = d A^T x + b y for the 4x4-block operator against scipy"""
    import scipy.sparse as sp
    lev = h8.fine
    A = H.CudaCSRMatrixInterface(ctx, lev.rowptr, lev.col, lev.val, 4)
    S = sp.bsr_matrix((lev.val.reshape(-1, 4, 4), lev.col, lev.rowptr),
                      shape=(4 * lev.n, 4 * lev.n))
    rng = np.random.default_rng(4)
    x, y0 = rng.uniform(-1, 1, 4 * lev.n), rng.uniform(-1, 1, 4 * lev.n)
    Y = H.CudaVectorInterface.from_global(ctx, y0, 4)
    A.vmult_transpose(Y, H.CudaVectorInterface.from_global(ctx, x, 4), -0.5, 2.0)
    want = -0.5 * (S.T @ x) + 2.0 * y0
    assert np.abs(Y.copy_back() - want).max() <= 1e-13 * np.abs(want).max()


@pytest.fixture(scope="module")
def h8full():
    return stokes_hierarchy(3, seed=5, with_full=True)


def test_flux_gradient_vs_oracle(ctx, h8full):
    """adjoint solve on the transposed levels (Vanka factors of the transposed
    patch matrices) and dJ/dg = (j - A^T z) on the inflow u_x entries: z
    against the oracle's adjoint to 1e-12, the gradient against the
    inflow-face masses that mass conservation fixes"""
    from gascoignelib_amd.autograd import StokesFluxProblem
    from test_stokes_adjoint_oracle import transposed
    print("[stokes flux gradient FP64 8^3]", flush=True)
    h = h8full
    P = StokesFluxProblem(ctx, h, prec="f64", krylov="gmres", tol=1e-12)
    grad, Z, rep = P.gradient()
    jF = P.j.copy()
    jF.reshape(-1, 4)[h.fine.dirichlet, 1:] = 0.0
    om = O.MG(transposed(h), smoother="vanka", patches=h.patches, **MGP)
    zo = np.zeros_like(jF)
    rc, orep = om.newton(zo, jF, O.cginfo(1e-12, 1e-30, 100), O.nlinfo(1e-12, 1e-30, 30),
                         norm_kind=1)
    assert rc == 0
    assert abs(rep.gmres_iterations - orep.gmres_iterations) <= 1
    assert np.linalg.norm(Z.copy_back() - zo) <= 1e-12 * np.linalg.norm(zo)
    want = h.mass_in[h.inflow_nodes]
    assert np.abs(grad - want).max() <= 1e-9 * want.max()
    # forward: outflow flux = inflow flux of the lifted data
    g = h.u0[h.inflow_nodes, 1]
    assert abs(P.J(g) - want @ g) <= 1e-9 * abs(want @ g)


def test_flux_gradient_mixed_and_autograd(ctx, h8full):
    """FP32 Vanka V-cycles in FP64 flexible GMRES for both solves; the torch
    wrapper's backward against central differences of forward solves"""
    import torch
    from gascoignelib_amd.autograd import InflowFlux, StokesFluxProblem
    print("[stokes flux gradient mixed + autograd 8^3]", flush=True)
    h = h8full
    P = StokesFluxProblem(ctx, h, prec="f32", krylov="fgmres", tol=1e-12)
    want = h.mass_in[h.inflow_nodes]
    g0 = h.u0[h.inflow_nodes, 1].copy()
    g = torch.tensor(g0, dtype=torch.float64, requires_grad=True)
    J = InflowFlux.apply(g, P)
    J.backward()
    assert np.abs(g.grad.numpy() - want).max() <= 1e-8 * want.max()
    assert abs(float(J.detach()) - want @ g0) <= 1e-9 * abs(want @ g0)
    rng = np.random.default_rng(8)
    eps = 1e-3
    for k in rng.choice(len(g0), 3, replace=False):
        gp, gm = g0.copy(), g0.copy()
        gp[k] += eps
        gm[k] -= eps
        fd = (P.J(gp) - P.J(gm)) / (2 * eps)
        assert abs(fd - float(g.grad[k])) <= 1e-6 * want.max()
