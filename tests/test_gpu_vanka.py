"""Vanka patch smoother on the GPU against the oracle.

FP64: the patch factors are formed with the oracle's operations (same pivots,
uncontracted updates) and the solves in the same order, so the smoother and the
V-cycle built on it agree bit for bit; GMRES / Newton iterates to 1e-12 with
iteration counts within one.  FP32 factors: within 1e-5 of FP64.
"""
import numpy as np
import pytest

import oracle_ctypes as O
from gascoignelib_amd import hip as H
from gascoignelib_amd.problem import Hierarchy, Level, block_operator, patches_from_cells
from gascoignelib_amd.structure import Mesh, assemble_hierarchy

pytestmark = pytest.mark.gpu

MGP = dict(mgtype="V", iterpre=2, iterpost=2, iterexact=4, omega=0.8, mgomega=1.0)


@pytest.fixture(scope="module")
def ctx():
    c = H.Context(0)
    yield c
    c.close()


def mesh_hierarchy(pre, B):
    """uniform cube hierarchy (levels 2^1 .. 2^pre cells per direction), scalar or
    the block operator of B components, and the Vanka patches of every level"""
    M = Mesh("cube.inp", 3, pre, 1)
    h = assemble_hierarchy(M)
    rng = np.random.default_rng(B)
    G = rng.uniform(-1, 1, (B, B))
    Cm, S = G @ G.T + B * np.eye(B), rng.uniform(-0.3, 0.3, (B, B))
    levels, patches = [], []
    for l, lev in enumerate(h.levels):
        val = lev.val if B == 1 else block_operator(lev, Cm, S)
        levels.append(Level(lev.n, B, lev.rowptr, lev.col, val, lev.transfer, lev.hn_kind,
                            dirichlet=lev.dirichlet,
                            dirichlet_comps=np.arange(B, dtype=np.int32), coords=lev.coords))
        patches.append(patches_from_cells(M.cells(l)))
    return Hierarchy(levels, 3, f"cube{pre}:B{B}"), patches


@pytest.mark.parametrize("B", [1, 4])
def test_vanka_solve_bitexact(ctx, B):
    h, patches = mesh_hierarchy(4 if B == 1 else 3, B)
    lev = h.fine
    A = H.CudaCSRMatrixInterface(ctx, lev.rowptr, lev.col, lev.val, B)
    x = np.random.default_rng(7).uniform(-1, 1, lev.n * B)
    X = H.CudaVectorInterface.from_global(ctx, x, B)
    H.VankaSmoother(ctx, A, patches[-1]).solve(X)
    xo = x.copy()
    O.Vanka(O.Matrix(lev.rowptr, lev.col, lev.val, B), patches[-1]).solve(xo)
    assert np.array_equal(X.copy_back(), xo)
    # FP32 factors and local solves
    X32 = H.CudaVectorInterface.from_global(ctx, x, B, "f32")
    H.VankaSmoother(ctx, A, patches[-1], "f32").solve(X32)
    assert np.abs(X32.copy_back() - xo).max() <= 1e-5 * np.abs(xo).max()


@pytest.mark.parametrize("B", [1, 4])
@pytest.mark.parametrize("order", [None, "lexicographic"])
def test_vcycle_with_vanka_vs_oracle(ctx, B, order):
    print(f"[vanka V-cycle B={B} order={order}]", flush=True)
    h, patches = mesh_hierarchy(4 if B == 1 else 3, B)
    f = np.random.default_rng(9).uniform(-1, 1, h.fine.n * B)
    for c in range(B):
        f[h.fine.dirichlet * B + c] = 0.0
    om = O.MG(h, ndirect=1, smoother="vanka", patches=patches, **MGP)
    xo = np.zeros_like(f)
    om.precondition(xo, f)
    mg = H.CudaMultiLevelSolver(ctx, h, ndirect=1, linear_smooth="vanka", patches=patches,
                                order=order, **MGP)
    X = mg.new_vector()
    mg.precondition(X, H.CudaVectorInterface.from_global(ctx, f, B))
    x = X.copy_back()
    if order is None:
        assert np.array_equal(x, xo)
    else:  # the same numbers, sums formed in the caller's order
        assert np.array_equal(x, xo)
    # Newton / GMRES with the Vanka V-cycle
    uo = np.zeros_like(f)
    _, orep = om.newton(uo, f, O.cginfo(1e-10, 1e-30, 100), O.nlinfo(1e-10, 1e-30, 30),
                        norm_kind=1)
    U = mg.new_vector()
    rep = mg.Solve(U, H.CudaVectorInterface.from_global(ctx, f, B), H.CGInfo(1e-10, 1e-30, 100),
                   H.NLInfo(1e-10, 1e-30, 30), norm_kind=1)
    u = U.copy_back()
    assert rep.newton_iterations == orep.newton_iterations
    assert abs(rep.gmres_iterations - orep.gmres_iterations) <= 1
    assert np.linalg.norm(u - uo) <= 1e-12 * np.linalg.norm(uo)


def test_vanka_mixed_precision_solve(ctx):
    h, patches = mesh_hierarchy(4, 4)
    f = np.random.default_rng(11).uniform(-1, 1, h.fine.n * 4)
    for c in range(4):
        f[h.fine.dirichlet * 4 + c] = 0.0
    om = O.MG(h, ndirect=1, smoother="vanka", patches=patches, **MGP)
    uo = np.zeros_like(f)
    om.newton(uo, f, O.cginfo(1e-12, 1e-30, 100), O.nlinfo(1e-12, 1e-30, 30), norm_kind=1)
    mg = H.CudaMultiLevelSolver(ctx, h, ndirect=1, linear_smooth="vanka", patches=patches,
                                prec="f32", krylov="fgmres", **MGP)
    U = mg.new_vector()
    rep = mg.Solve(U, H.CudaVectorInterface.from_global(ctx, f, 4), H.CGInfo(0.9e-10, 1e-30, 100),
                   H.NLInfo(1e-10, 1e-30, 30), norm_kind=1)
    assert H.N.NL_STATUS[rep.status] == "converged"
    assert np.linalg.norm(U.copy_back() - uo) <= 1e-8 * np.linalg.norm(uo)
