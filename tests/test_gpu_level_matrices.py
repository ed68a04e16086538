"""Every level matrix of the ordered solver against the oracle, on matrices
whose rows follow many column patterns (the variable-coefficient Q1 operator
at 96^3 cells and adaptive Q2 with hanging nodes), where the device layout
groups rows by length and pattern, runs slices from the pattern table and
keeps others in 16- or 32-bit columns.

* FP64 solver, forward and adjoint: in the solver's internal numbering, with
  each row's entries in the order the solver keeps (the caller's order; the
  transposed rows of the adjoint list their columns ascending in the caller's
  numbering), the residual b - A x equals the oracle's bit for bit.
* Reduced-precision V-cycle: the f16-value matrix agrees with the FP64 one to
  f16 rounding of the values.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle_ctypes as O
from gascoignelib_amd import hip as H
from gascoignelib_amd.problem import Level, lexicographic_order, permuted_csr
from gascoignelib_amd.structure import q2_hierarchy

pytestmark = pytest.mark.gpu

MGP = dict(mgtype="V", iterpre=2, iterpost=2, iterexact=10, omega=0.8, mgomega=1.0,
           order="lexicographic")


@pytest.fixture(scope="module")
def ctx():
    c = H.Context(0)
    yield c
    c.close()


def diffusion96(ctx):
    from gascoignelib_amd.autograd import DiffusionProblem
    from gascoignelib_amd.problem import cell_centres, gaussian_random_field
    P = DiffusionProblem(ctx, inp="cube3.inp", prerefine=5, ndirect=400)
    k = np.exp(gaussian_random_field(cell_centres(P.coords, P.cells), seed=31))
    return P.hierarchy(k), 400


def q2_adaptive(ctx):
    return q2_hierarchy(4, adaptive=2, radius=0.25, centre=(0.5, 0.47, 0.53)), 200


def transposed(lev):
    T = sp.csr_matrix((lev.val, lev.col, lev.rowptr), shape=(lev.n, lev.n)).T.tocsr()
    T.sort_indices()
    return Level(lev.n, 1, T.indptr.astype(np.int32), T.indices.astype(np.int32),
                 np.ascontiguousarray(T.data), coords=lev.coords)


@pytest.mark.parametrize("adjoint", [False, True])
@pytest.mark.parametrize("make", [diffusion96, q2_adaptive], ids=["diffusion96", "q2ada16"])
def test_fp64_level_matrices_bitexact(ctx, make, adjoint):
    h, ndirect = make(ctx)
    print(f"[{make.__name__} adjoint={adjoint}: FP64 level matrices]", flush=True)
    mg = H.CudaMultiLevelSolver(ctx, h, adjoint=adjoint, ndirect=ndirect, **MGP)
    rng = np.random.default_rng(3)
    checked = 0
    for li, lev in enumerate(h.levels):
        if lev.n < ndirect:
            continue  # direct levels keep no SELL matrix of interest
        order = lexicographic_order(lev.coords)
        src = transposed(lev) if adjoint else lev
        rp, col, val = permuted_csr(src, order, sort_columns=False)
        A = mg.level_matrix(li, fp64=True)
        assert A.n() == lev.n
        x, b = rng.uniform(-1, 1, lev.n), rng.uniform(-1, 1, lev.n)
        R = H.CudaVectorInterface(ctx, lev.n)
        A.residual(R, H.CudaVectorInterface.from_global(ctx, x),
                   H.CudaVectorInterface.from_global(ctx, b))
        r = np.zeros(lev.n)
        O.Matrix(rp, col, val).residual(r, x, b)
        assert np.array_equal(R.copy_back(), r), f"level {li}"
        checked += 1
        del A
    mg.free()
    assert checked >= 2


@pytest.mark.parametrize("make", [diffusion96, q2_adaptive], ids=["diffusion96", "q2ada16"])
def test_f16_level_matrices(ctx, make):
    h, ndirect = make(ctx)
    print(f"[{make.__name__}: f16 level matrices]", flush=True)
    mg = H.CudaMultiLevelSolver(ctx, h, prec="f16", krylov="fgmres", ndirect=ndirect, **MGP)
    rng = np.random.default_rng(4)
    for li, lev in enumerate(h.levels):
        if lev.n < ndirect:
            continue
        order = lexicographic_order(lev.coords)
        rp, col, val = permuted_csr(lev, order, sort_columns=True)
        M = sp.csr_matrix((val, col, rp), shape=(lev.n, lev.n))
        A = mg.level_matrix(li, fp64=False)
        x = rng.uniform(-1, 1, lev.n)
        Y = H.CudaVectorInterface(ctx, lev.n, 1, "f32")
        A.vmult(Y, H.CudaVectorInterface.from_global(ctx, x, 1, "f32"))
        want = M @ x.astype(np.float32).astype(np.float64)
        # f16 values (relative rounding 2^-11, absolute 2^-25 below the
        # normal range) summed in FP32
        nz = (M != 0).astype(np.float64)
        bound = (2.0 ** -10 * (abs(M) @ np.abs(x)) + 2.0 ** -24 * (nz @ np.abs(x))
                 + 1e-6 * np.abs(want).max())
        assert np.all(np.abs(Y.copy_back() - want) <= bound), f"level {li}"
        del A
    mg.free()
