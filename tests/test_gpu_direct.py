"""Direct coarse solve on the device (the dense inverse factorised by
rocSOLVER getrf / getri, applied by k_dense_solve).

With ndirect above the finest level's size the whole multigrid is one direct
solve, so the preconditioner applied to b is A^-1 b: its backward error
must be at rounding level, for a scalar 17^3 operator (4,913
unknowns, the largest level a direct solve covers here) and for the 4x4
Stokes-LPS blocks (2,916 unknowns), in FP64 and with the FP32 V-cycle
vectors, measured as the normwise backward error.  A singular level is an
error."""
import numpy as np
import pytest
import scipy.sparse as sp

from gascoignelib_amd import hip as H
from gascoignelib_amd import native as N
from gascoignelib_amd.problem import Hierarchy, Level, splitmix64_uniform
from gascoignelib_amd.structure import stokes_hierarchy, uniform_hierarchy

pytestmark = pytest.mark.gpu

MGP = dict(mgtype="V", iterpre=2, iterpost=2, iterexact=4, omega=0.8, mgomega=1.0)


@pytest.fixture(scope="module")
def ctx():
    c = H.Context(0)
    yield c
    c.close()


def dense_of(lev):
    B = lev.ncomp
    V = lev.val.reshape(-1, B, B)
    A = sp.bsr_matrix((V, lev.col, lev.rowptr), shape=(lev.n * B, lev.n * B)).tocsr()
    return A


@pytest.mark.parametrize("case", ["q1_17", "stokes_8"])
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_whole_hierarchy_direct_is_exact(ctx, case, prec):
    print(f"[direct {case} {prec}]", flush=True)
    h = uniform_hierarchy("cube.inp", 3, 4, drop_coarse=1) if case == "q1_17" else \
        stokes_hierarchy(3, drop_coarse=1)
    fine = h.fine
    m = fine.n * fine.ncomp
    mg = H.CudaMultiLevelSolver(ctx, h, ndirect=fine.n + 1, prec=prec, **MGP)
    b = splitmix64_uniform(7, m)
    X = mg.new_vector()
    mg.precondition(X, H.CudaVectorInterface.from_global(ctx, b, fine.ncomp), x_is_zero=True)
    x = X.copy_back()
    A = dense_of(fine)
    # normwise backward error |b - A x| / (|A| |x| + |b|) (infinity norms): the
    # Stokes saddle point has |A| |x| >> |b|, so the plain relative residual
    # reflects its conditioning, not the solve
    r = np.abs(A @ x - b).max()
    eta = r / (abs(A).sum(axis=1).max() * np.abs(x).max() + np.abs(b).max())
    assert eta <= (1e-14 if prec == "f64" else 1e-6), eta


def test_singular_direct_level_is_an_error(ctx):
    print("[direct singular]", flush=True)
    # [[1, 1], [1, 1]]: nonzero diagonal (the level matrix is accepted), the
    # second LU pivot is exactly zero
    lev = Level(2, 1, np.array([0, 2, 4]), np.array([0, 1, 0, 1]), np.ones(4), None, 13,
                dirichlet=np.zeros(0, np.int32), coords=np.zeros((2, 3)))
    h = Hierarchy([lev], 3, "singular")
    with pytest.raises(RuntimeError, match="singular"):
        H.CudaMultiLevelSolver(ctx, h, ndirect=3, **MGP)
