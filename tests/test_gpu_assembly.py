"""Device assembly (gh_assemble_q1, gh_form_q1, gh_mg_assemble_q1) on the GPU.

The reference assembles on the CPU (CGDisc::Matrix, src/DofHandler/cgdisc.h:
599-635, with SparseBlockMatrix::dirichlet, src/LinAlg/sparseblockmatrix.xx:
205-218) and its GPU path uploads the result on every assembly
(cuda/cudasolver.cc:114-128).  Checks:

* on Gascoigne's own meshes (compiled-reference fixtures), the device values
  equal the reference's assembled values to rounding (1e-13 relative, the
  bar the host producer meets) and the host producer's values bit for bit;
* with a per-cell coefficient, bit for bit against the host producer, in
  FP64, FP32 and FP16 storage, including the inverse diagonal (a Jacobi
  application) and the device layout (products);
* the cell-by-cell residual f - A(u) against the assembled residual;
* misuse is an error: a pattern that lacks a coupling, and an entry left out
  of the layout (an exact zero at creation) that the new cells make nonzero;
* re-assembling a whole solver on the device with a new coefficient gives the
  solves of a solver built from host-assembled matrices, bit for bit.
"""
import numpy as np
import pytest

from gascoignelib_amd import hip as H
from gascoignelib_amd import native as N
from gascoignelib_amd.problem import lattice_rhs
from gascoignelib_amd.structure import Mesh, assemble_hierarchy, coarse_coefficients

pytestmark = pytest.mark.gpu

FIXTURES = ["q1_2d_128", "q1_3d_16", "q1_3d_cube3_12"]


@pytest.fixture(scope="module")
def ctx():
    c = H.Context(0)
    yield c
    c.close()


def host_assemble(dim, coords, cells, rowptr, col, dirichlet, kcell=None):
    xyz = np.zeros((len(coords), 3))
    xyz[:, :dim] = coords
    cells = np.ascontiguousarray(cells, dtype=np.int32)
    d = np.ascontiguousarray(dirichlet, dtype=np.int32)
    val = np.empty(len(col))
    kp = None
    if kcell is not None:
        kcell = np.ascontiguousarray(kcell, dtype=np.float64)
        kp = N.ptr(kcell)
    N.check_s(N.structure().ghs_assemble_on(dim, len(xyz), N.ptr(xyz), len(cells),
                                            N.ptr(cells, N.i32p), kp,
                                            N.ptr(np.ascontiguousarray(rowptr), N.i32p),
                                            N.ptr(np.ascontiguousarray(col), N.i32p), len(d),
                                            N.ptr(d, N.i32p), N.ptr(val)))
    return val


@pytest.mark.parametrize("name", FIXTURES)
def test_device_assembly_on_reference_meshes(ctx, golden, name):
    print(f"[device assembly {name}]", flush=True)
    z = golden(name)
    dim = int(z["dim"])
    for l in range(int(z["nlevels"])):
        p = f"L{l}_"
        if p + "val" not in z.files:
            continue
        rp, col, ref = z[p + "rowptr"], z[p + "col"], z[p + "val"]
        n = len(rp) - 1
        coords, cells, bnd = z[p + "coords"], z[p + "cells"], z[p + "bnd1"]
        host = host_assemble(dim, coords, cells, rp, col, bnd)
        # created with zeros in place of the values: every entry stored, so
        # the device layout holds each coupling the assembly writes
        A = H.CudaCSRMatrixInterface(ctx, rp, col, np.where(ref != 0.0, ref, 1.0))
        Cc = H.CoefficientCells(ctx, dim, coords, cells)
        M = H.DirichletMask(ctx, n, 1, bnd)
        Cc.assemble(A, mask=M)
        got = A.values()
        assert np.array_equal(got, host), f"level {l}: device differs from the host producer"
        assert np.allclose(got, ref, rtol=1e-13, atol=1e-15 * np.abs(ref).max())
        # the product runs on the refreshed device layout
        x = np.random.default_rng(l).uniform(-1, 1, n)
        X, Y = H.CudaVectorInterface.from_global(ctx, x), H.CudaVectorInterface(ctx, n)
        A.vmult(Y, X)
        import scipy.sparse as sp
        yh = sp.csr_matrix((host, col, rp), shape=(n, n)) @ x
        np.testing.assert_allclose(Y.copy_back(), yh, rtol=1e-13, atol=1e-13 * np.abs(yh).max())


def cube(pre):
    M = Mesh("cube.inp", 3, pre, 1)
    return M


@pytest.mark.parametrize("prec", ["f64", "f32", "f16"])
def test_device_assembly_variable_coefficient_bitexact(ctx, prec):
    print(f"[device assembly k(x) {prec}]", flush=True)
    M = cube(4)
    l = M.nlevels - 1
    coords, cells = M.coords(l), M.cells(l)
    rp, col = M.structure(l)
    n = len(rp) - 1
    dirichlet = M.dirichlet(l)
    rng = np.random.default_rng(11)
    k = np.exp(rng.normal(size=len(cells)))
    host1 = host_assemble(3, coords, cells, rp, col, dirichlet)           # k = 1
    host2 = host_assemble(3, coords, cells, rp, col, dirichlet, k)        # k(x)
    # created from the k = 1 values, re-assembled with k(x) on the device
    A = H.CudaCSRMatrixInterface(ctx, rp, col, host1, 1, prec)
    Ref = H.CudaCSRMatrixInterface(ctx, rp, col, host2, 1, prec)
    Cc = H.CoefficientCells(ctx, 3, coords, cells)
    K = H.CudaVectorInterface.from_global(ctx, k)
    Cc.assemble(A, k=K, mask=H.DirichletMask(ctx, n, 1, dirichlet))
    assert np.array_equal(A.values(), Ref.values())
    # layout values and inverse diagonal: products, residuals and Jacobi equal
    x = rng.uniform(-1, 1, n)
    vp = "f64" if prec == "f64" else "f32"
    outs = []
    for Mx in (A, Ref):
        X = H.CudaVectorInterface.from_global(ctx, x, 1, vp)
        Y = H.CudaVectorInterface(ctx, n, 1, vp)
        Mx.vmult(Y, X)
        Mx.jacobi(X)
        outs.append((Y.copy_back(), X.copy_back()))
    assert np.array_equal(outs[0][0], outs[1][0])
    assert np.array_equal(outs[0][1], outs[1][1])


def test_form_matches_assembled_residual(ctx):
    print("[form q1]", flush=True)
    for dim, inp, pre in ((3, "cube.inp", 4), (2, "square.inp", 5)):
        M = Mesh(inp, dim, pre, 1)
        l = M.nlevels - 1
        coords, cells = M.coords(l), M.cells(l)
        rp, col = M.structure(l)
        n = len(rp) - 1
        d = M.dirichlet(l)
        rng = np.random.default_rng(dim)
        k = np.exp(rng.normal(size=len(cells)))
        val = host_assemble(dim, coords, cells, rp, col, d, k)
        u = rng.uniform(-1, 1, n)
        u[d] = 0.0
        f = rng.uniform(-1, 1, n)
        Cc = H.CoefficientCells(ctx, dim, coords, cells)
        Mk = H.DirichletMask(ctx, n, 1, d)
        K = H.CudaVectorInterface.from_global(ctx, k)
        U, F, R = (H.CudaVectorInterface.from_global(ctx, v) for v in (u, f, np.zeros(n)))
        Cc.form(R, U, F, k=K, mask=Mk)
        import scipy.sparse as sp
        r = f - sp.csr_matrix((val, col, rp), shape=(n, n)) @ u
        r[d] = 0.0
        got = R.copy_back()
        np.testing.assert_allclose(got, r, rtol=0, atol=1e-13 * np.abs(f).max())
        assert np.all(got[d] == 0.0)


def test_assembly_errors(ctx):
    print("[device assembly errors]", flush=True)
    lib = N.hip()
    M = cube(3)
    l = M.nlevels - 1
    coords, cells = M.coords(l), M.cells(l)
    rp, col = M.structure(l)
    n = len(rp) - 1
    # a diagonal-only pattern misses the couplings of the cells
    Dg = H.CudaCSRMatrixInterface(ctx, np.arange(n + 1), np.arange(n), np.ones(n))
    Cc = H.CoefficientCells(ctx, 3, coords, cells)
    rc = lib.gh_assemble_q1(ctx.h, Cc.h, None, None, Dg.h)
    assert rc == N.GH_EINVAL and "pattern" in lib.gh_last_error().decode()
    # the unit cube's face couplings are exact zeros (left out of the layout);
    # cells stretched in x make them nonzero
    val = host_assemble(3, coords, cells, rp, col, M.dirichlet(l))
    assert np.count_nonzero(val == 0.0) > 0
    A = H.CudaCSRMatrixInterface(ctx, rp, col, val)
    Cs = H.CoefficientCells(ctx, 3, coords * np.array([2.0, 1.0, 1.0]), cells)
    rc = lib.gh_assemble_q1(ctx.h, Cs.h, None, None, A.h)
    assert rc == N.GH_EINVAL and "zero at creation" in lib.gh_last_error().decode()
    # a coefficient vector of the wrong size
    K = H.CudaVectorInterface(ctx, len(cells) + 1)
    assert lib.gh_assemble_q1(ctx.h, Cc.h, K.h, None, A.h) == N.GH_EINVAL
    # the refused calls left the matrix usable: a proper assembly still works
    Cc.assemble(A, mask=H.DirichletMask(ctx, n, 1, M.dirichlet(l)))
    assert np.array_equal(A.values(), val)


def level_cells(hier):
    M = hier.mesh
    return [(M.coords(l), M.cells(l)) for l in range(M.nlevels)]


@pytest.mark.parametrize("prec,krylov", [("f64", "gmres"), ("f16", "fgmres")])
def test_solver_reassembly_matches_host_built_solver(ctx, prec, krylov):
    print(f"[mg re-assembly {prec} {krylov}]", flush=True)
    M = cube(4)
    ncf = M.sizes(M.nlevels - 1)[1]
    rng = np.random.default_rng(5)
    k1 = np.exp(rng.normal(size=ncf))
    k2 = np.exp(0.5 * rng.normal(size=ncf))
    h1 = assemble_hierarchy(M, kcell=k1)
    h2 = assemble_hierarchy(M, kcell=k2)
    params = dict(mgtype="V", iterpre=2, iterpost=2, iterexact=10, omega=0.8, mgomega=1.0,
                  ndirect=200, prec=prec, krylov=krylov, order="lexicographic")
    mg1 = H.CudaMultiLevelSolver(ctx, h1, **params)
    mg2 = H.CudaMultiLevelSolver(ctx, h2, **params)
    cells = [H.CoefficientCells(ctx, 3, c, cl) for c, cl in level_cells(h1)]
    kv = [H.CudaVectorInterface.from_global(ctx, k) for k in h2.kcell]
    mg1.assemble_q1(cells, kv)
    f = lattice_rhs(h2.fine.coords, seed=3)
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
    # the level matrices the two solvers run on hold the same values
    for l in range(h1.nlevels):
        for fp64 in (True, False):
            assert np.array_equal(mg1.level_matrix(l, fp64).values(),
                                  mg2.level_matrix(l, fp64).values())
