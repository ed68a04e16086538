"""Block-Jacobi smoothing (B = 2..8) on the GPU against the oracle.

The reference's one hand-written CUDA kernel inverts the diagonal blocks
(cuda/cusparsehelper.cu:24-118); the CPU path it must agree with inverts each
block by Gauss-Jordan with partial pivoting (src/LinAlg/nodematrix.h:159-213)
inside every Jacobi application (SparseBlockMatrix::Jacobi,
src/LinAlg/sparseblockmatrix.xx:47-56; FMatrixBlock::inverse / vmult,
src/LinAlg/fmatrixblock.xx:247-252).  FP64 results are compared bit for bit:
h = D^-1 h, smoothing sweeps x += omega D^-1 (b - A x), and whole V-cycles
with block-Jacobi smoothing (component-row layout for B = 2..4, node-block
CSR for B = 5..8).  In the product tests the diagonal blocks are
non-symmetric and need row exchanges, so the pivot order is exercised.
"""
import numpy as np
import pytest

import oracle_ctypes as O
from gascoignelib_amd import hip as H
from gascoignelib_amd.problem import Hierarchy, Level, block_operator
from gascoignelib_amd.structure import Mesh, assemble_hierarchy

pytestmark = pytest.mark.gpu

MGP = dict(mgtype="V", iterpre=2, iterpost=2, iterexact=4, omega=0.8, mgomega=1.0)


@pytest.fixture(scope="module")
def ctx():
    c = H.Context(0)
    yield c
    c.close()


def block_blocks(B, seed, pivot):
    """coupling C (symmetric positive definite, so the operator stays
    elliptic) and a diagonal shift S.  pivot: the last row of S gets a first
    entry of 2B + 1, above the diagonal block's (0, 0) entry on a grid of
    spacing 1/8 (at most 8h/3 * 2B + 0.3), so every interior diagonal block
    needs a row exchange in the first Gauss-Jordan step"""
    rng = np.random.default_rng(seed)
    G = rng.uniform(-1, 1, (B, B))
    Cm = G @ G.T + B * np.eye(B)
    S = rng.uniform(-0.3, 0.3, (B, B))
    if pivot:
        S[B - 1, 0] = 2.0 * B + 1.0
    return Cm, S


def block_hierarchy(pre, B, seed, pivot=False):
    M = Mesh("cube.inp", 3, pre, 1)
    h = assemble_hierarchy(M)
    Cm, S = block_blocks(B, seed, pivot)
    levels = []
    for lev in h.levels:
        levels.append(Level(lev.n, B, lev.rowptr, lev.col, block_operator(lev, Cm, S),
                            lev.transfer, lev.hn_kind, dirichlet=lev.dirichlet,
                            dirichlet_comps=np.arange(B, dtype=np.int32), coords=lev.coords))
    return Hierarchy(levels, 3, f"cube{pre}:B{B}")


def pivots_needed(val, rowptr, col, B):
    """True when some diagonal block's first elimination step exchanges rows"""
    rows = np.repeat(np.arange(len(rowptr) - 1), np.diff(rowptr))
    D = val.reshape(-1, B, B)[rows == col]
    return bool(np.any(np.argmax(np.abs(D[:, :, 0]), axis=1) > 0))


@pytest.mark.parametrize("B", [2, 3, 4, 5, 6, 7, 8])
def test_block_jacobi_and_sweeps_bitexact(ctx, B):
    print(f"[block jacobi B={B}]", flush=True)
    lev = block_hierarchy(3, B, B, pivot=True).fine
    assert pivots_needed(lev.val, lev.rowptr, lev.col, B)
    A = H.CudaCSRMatrixInterface(ctx, lev.rowptr, lev.col, lev.val, B)
    Ao = O.Matrix(lev.rowptr, lev.col, lev.val, B)
    rng = np.random.default_rng(100 + B)
    m = lev.n * B
    h, x, b = (rng.uniform(-1, 1, m) for _ in range(3))
    # h = D^-1 h
    Hv = H.CudaVectorInterface.from_global(ctx, h, B)
    A.jacobi(Hv)
    ho = h.copy()
    Ao.jacobi(ho)
    assert np.array_equal(Hv.copy_back(), ho)
    # D^-1 D = I on a few blocks (the oracle's Gauss-Jordan against numpy)
    rows = np.repeat(np.arange(lev.n), np.diff(lev.rowptr))
    D = lev.val.reshape(-1, B, B)[rows == lev.col]
    for i in (0, lev.n // 2, lev.n - 1):
        e = np.zeros(m)
        e[i * B:(i + 1) * B] = rng.uniform(-1, 1, B)
        E = H.CudaVectorInterface.from_global(ctx, e, B)
        A.jacobi(E)
        np.testing.assert_allclose(D[i] @ E.copy_back()[i * B:(i + 1) * B], e[i * B:(i + 1) * B],
                                   rtol=0, atol=1e-13)
    # three sweeps x += 0.8 D^-1 (b - A x): StdSolver::smooth, jacobi branch
    X = H.CudaVectorInterface.from_global(ctx, x, B)
    Tmp = H.CudaVectorInterface(ctx, lev.n, B)
    A.smooth(X, H.CudaVectorInterface.from_global(ctx, b, B), Tmp, 3, 0.8)
    xo, r = x.copy(), np.zeros(m)
    for _ in range(3):
        Ao.residual(r, xo, b)
        Ao.jacobi(r)
        xo += 0.8 * r
    assert np.array_equal(X.copy_back(), xo)


@pytest.mark.parametrize("B", [2, 3, 4, 5, 8])
@pytest.mark.parametrize("mgtype", ["V", "W"])
def test_block_jacobi_vcycle_bitexact(ctx, B, mgtype):
    print(f"[block jacobi V-cycle B={B} {mgtype}]", flush=True)
    h = block_hierarchy(3, B, 50 + B)
    f = np.random.default_rng(9).uniform(-1, 1, h.fine.n * B)
    for c in range(B):
        f[h.fine.dirichlet * B + c] = 0.0
    params = dict(MGP, mgtype=mgtype)
    om = O.MG(h, ndirect=1, smoother="jacobi", **params)
    xo = np.zeros_like(f)
    om.precondition(xo, f)
    mg = H.CudaMultiLevelSolver(ctx, h, ndirect=1, linear_smooth="jacobi", **params)
    X = mg.new_vector()
    mg.precondition(X, H.CudaVectorInterface.from_global(ctx, f, B))
    assert np.array_equal(X.copy_back(), xo)
    assert mg.mgsteps == om.mgsteps
    # the preconditioned solve converges and matches the oracle's GMRES
    uo = np.zeros_like(f)
    io = O.cginfo(1e-10, 1e-30, 60)
    om.gmres(uo, f, io, vmax=30)
    U = mg.new_vector()
    info = H.CGInfo(1e-10, 1e-30, 60)
    mg.Gmres(U, H.CudaVectorInterface.from_global(ctx, f, B), info, vmax=30)
    assert info.status == O.CG_STATUS[io.status] == "converged"
    assert info.iteration == io.iteration
    u = U.copy_back()
    assert np.linalg.norm(u - uo) <= 1e-12 * np.linalg.norm(uo)
