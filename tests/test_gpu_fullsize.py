"""The workloads at their full sizes, on the GPU.

At these sizes the oracle still runs single operator applications and one
V-cycle in seconds, so those are compared with it bit for bit; whole solves
are checked through properties that do not depend on the size: the true FP64
residual, symmetry of the operator and of the V(2,2) Jacobi preconditioner,
the mixed solution against the FP64 one, mass conservation and the closed-form
flux gradient of the Stokes channel, Euler's identity for the coefficient
gradient, and the duality g.u = z.f of forward and adjoint solves.

Sizes: 3D Q1 Poisson 128^3 cells (2,146,689 DOF); Stokes-LPS channel 64^3
cells (1,098,500 unknowns); variable-coefficient diffusion 96^3 cells
(912,673 DOF); Q2 patch-element Poisson 64^3 cells (274,625 DOF), and the
same with two adaptive levels and Q2 hanging nodes (1,986,413 DOF).
"""
import numpy as np
import pytest

import oracle_ctypes as O
from gascoignelib_amd import hip as H
from gascoignelib_amd.problem import lattice_rhs, splitmix64_uniform
from gascoignelib_amd.structure import q2_hierarchy, stokes_hierarchy, uniform_hierarchy

pytestmark = pytest.mark.gpu

MGP = dict(mgtype="V", iterpre=2, iterpost=2, iterexact=10, omega=0.8, mgomega=1.0)


@pytest.fixture(scope="module")
def ctx():
    c = H.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def p128():
    h = uniform_hierarchy("cube.inp", 3, 7, drop_coarse=1)
    f = lattice_rhs(h.fine.coords, seed=2024)
    f[h.fine.dirichlet] = 0.0
    return h, f


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_poisson128_operators_and_vcycle_vs_oracle(ctx, p128):
    print("[128^3 operators and V-cycle vs oracle]", flush=True)
    h, f = p128
    lev = h.fine
    x = splitmix64_uniform(5, lev.n)
    A = H.CudaCSRMatrixInterface(ctx, lev.rowptr, lev.col, lev.val)
    Ao = O.Matrix(lev.rowptr, lev.col, lev.val)
    X, F = H.CudaVectorInterface.from_global(ctx, x), H.CudaVectorInterface.from_global(ctx, f)
    R = H.CudaVectorInterface(ctx, lev.n)
    A.residual(R, X, F)
    ro = np.zeros(lev.n)
    Ao.residual(ro, x, f)
    assert np.array_equal(R.copy_back(), ro)
    # one V(2,2) cycle, iterative coarse solve: bit for bit
    om = O.MG(h, ndirect=1, **MGP)
    xo = np.zeros(lev.n)
    om.precondition(xo, f)
    mg = H.CudaMultiLevelSolver(ctx, h, ndirect=1, order="lexicographic", **MGP)
    Y = mg.new_vector()
    mg.precondition(Y, F)
    assert np.array_equal(Y.copy_back(), xo)


def test_poisson128_solves(ctx, p128):
    print("[128^3 FP64 and mixed solves]", flush=True)
    h, f = p128
    lev = h.fine
    common = dict(MGP, ndirect=1000, order="lexicographic")
    mg = H.CudaMultiLevelSolver(ctx, h, **common)
    F = H.CudaVectorInterface.from_global(ctx, f)
    A = mg.fine_matrix
    # the operator and the V(2,2) Jacobi preconditioner are symmetric
    rng = np.random.default_rng(6)
    x, y = rng.uniform(-1, 1, lev.n), rng.uniform(-1, 1, lev.n)
    x[lev.dirichlet] = y[lev.dirichlet] = 0.0
    X, Y = H.CudaVectorInterface.from_global(ctx, x), H.CudaVectorInterface.from_global(ctx, y)
    AX, AY = H.CudaVectorInterface(ctx, lev.n), H.CudaVectorInterface(ctx, lev.n)
    A.vmult(AX, X)
    A.vmult(AY, Y)
    assert abs(AX.dot(Y) - X.dot(AY)) <= 1e-13 * abs(AX.dot(Y))
    MX, MY = mg.new_vector(), mg.new_vector()
    mg.precondition(MX, X, x_is_zero=True)
    mg.precondition(MY, Y, x_is_zero=True)
    assert abs(MX.dot(Y) - X.dot(MY)) <= 1e-12 * abs(MX.dot(Y))

    def true_res(U):
        R = H.CudaVectorInterface(ctx, lev.n)
        A.residual(R, U, F)
        mg.fine_mask.zero(R)
        return R.norm() / F.norm()

    U = mg.new_vector()
    rep = mg.Solve(U, F, H.CGInfo(1e-10, 1e-30, 100), H.NLInfo(1e-10, 1e-30, 50), norm_kind=1)
    assert H.N.NL_STATUS[rep.status] == "converged" and rep.gmres_iterations == 8
    assert true_res(U) <= 1e-10
    u = U.copy_back()
    mx = H.CudaMultiLevelSolver(ctx, h, prec="f16", krylov="fgmres",
                                **dict(common, iterpre=3, iterpost=2))
    V = mx.new_vector()
    rep = mx.Solve(V, F, H.CGInfo(0.9e-10, 1e-30, 100), H.NLInfo(1e-10, 1e-30, 50),
                   norm_kind=1)
    assert H.N.NL_STATUS[rep.status] == "converged"
    assert true_res(V) <= 1e-10
    assert rel(V.copy_back(), u) <= 1e-8


def test_stokes64_flux_and_gradient(ctx):
    """forward and flux-gradient solves of the 64^3 channel in FP64 and mixed
    precision: outflow flux = inflow flux, gradient = inflow-face masses"""
    from gascoignelib_amd.stokes import StokesFluxProblem
    print("[stokes 64^3 flux and gradient]", flush=True)
    h = stokes_hierarchy(6, seed=2024, with_full=True)
    g = h.u0[h.inflow_nodes, 1].copy()
    want = h.mass_in[h.inflow_nodes]
    fin = float(want @ g)
    sol = {}
    for prec, kry in (("f64", "gmres"), ("f32", "fgmres")):
        P = StokesFluxProblem(ctx, h, prec=prec, krylov=kry, tol=1e-10)
        u, rep = P.state(g)
        assert H.N.NL_STATUS[rep.status] == "converged"
        assert abs(P.j @ u - fin) <= 1e-9 * fin
        grad, _, _ = P.gradient()
        assert np.abs(grad - want).max() <= 1e-8 * want.max()
        sol[prec] = u
        del P
    assert rel(sol["f32"], sol["f64"]) <= 1e-8


def test_diffusion96_gradient_identity(ctx):
    """96^3 variable coefficient: sum_c k_c dJ/dk_c = -J (J = g.u is
    homogeneous of degree -1 in k), FP64"""
    from gascoignelib_amd.autograd import DiffusionProblem
    from gascoignelib_amd.problem import cell_centres, gaussian_random_field
    print("[diffusion 96^3 Euler identity]", flush=True)
    P = DiffusionProblem(ctx, inp="cube3.inp", prerefine=5, ndirect=400, tol=1e-11)
    k = np.exp(gaussian_random_field(cell_centres(P.coords, P.cells), seed=31))
    hier, U, _ = P.state(k)
    J = float(P.g @ U.copy_back())
    _, dk = P.gradient(hier, U)
    assert abs(k @ dk + J) <= 1e-9 * abs(J)


def test_q2_64_duality(ctx):
    """Q2 64^3: forward A u = f and adjoint A^T z = g give g.u = z.f"""
    print("[q2 64^3 duality]", flush=True)
    h = q2_hierarchy(6)
    lev = h.fine
    f = lattice_rhs(lev.coords, seed=2024)
    f[lev.dirichlet] = 0.0
    g = splitmix64_uniform(77, lev.n)
    g[lev.dirichlet] = 0.0
    common = dict(MGP, ndirect=1000, order="lexicographic")
    out = []
    for adjoint, rhs in ((False, f), (True, g)):
        mg = H.CudaMultiLevelSolver(ctx, h, adjoint=adjoint, **common)
        X = mg.new_vector()
        rep = mg.Solve(X, H.CudaVectorInterface.from_global(ctx, rhs),
                       H.CGInfo(1e-12, 1e-30, 200), H.NLInfo(1e-12, 1e-30, 50), norm_kind=1)
        assert H.N.NL_STATUS[rep.status] == "converged"
        out.append(X.copy_back())
        mg.free()
    u, z = out
    assert abs(g @ u - z @ f) <= 1e-10 * abs(g @ u)


def test_q2_adaptive_64_solves(ctx):
    """Q2 64^3 cells + 2 adaptive levels in the ball of radius 0.25 around the
    seeded centre (1,986,413 DOF, 120.8 M entries, 63,008 edge and 31,504 face
    hanging nodes, the reference's refinement and levels): one V-cycle bit for bit against the oracle; FP64 and mixed
    solves to the true FP64 residual 1e-10, mixed within 1e-8 of FP64; forward
    and adjoint solves satisfy g.u = z.f"""
    print("[q2 64^3 + 2 adaptive levels]", flush=True)
    import scipy.sparse as sp
    centre = 0.25 + 0.5 * (0.5 * (splitmix64_uniform(4242, 3) + 1.0))
    h = q2_hierarchy(6, adaptive=2, radius=0.25, centre=tuple(centre))
    lev = h.fine
    assert lev.n == 1986413 and len(lev.hn_edges) == 63008 and len(lev.hn_faces) == 31504
    hn = O.HN(lev.hn_kind, lev.hn_edges, lev.hn_faces)
    f = lattice_rhs(lev.coords, seed=2024)
    hn.distribute(f)
    f[lev.dirichlet] = 0.0
    g = (np.linalg.norm(lev.coords - centre, axis=1) < 0.25).astype(float)
    hn.distribute(g)
    g[lev.dirichlet] = 0.0
    om = O.MG(h, ndirect=1, **MGP)
    xo = np.zeros(lev.n)
    om.precondition(xo, f)
    mg = H.CudaMultiLevelSolver(ctx, h, ndirect=1, order="lexicographic", **MGP)
    Y = mg.new_vector()
    mg.precondition(Y, H.CudaVectorInterface.from_global(ctx, f))
    assert np.array_equal(Y.copy_back(), xo)
    mg.free()
    del om
    A = sp.csr_matrix((lev.val, lev.col, lev.rowptr), shape=(lev.n, lev.n))
    common = dict(MGP, ndirect=1000, order="lexicographic")
    F = H.CudaVectorInterface.from_global(ctx, f)
    out = {}
    for tag, kw, tol in (("fp64", {}, 1e-10), ("mixed", dict(prec="f16", krylov="fgmres"), 0.9e-10)):
        mg = H.CudaMultiLevelSolver(ctx, h, **kw, **common)
        U = mg.new_vector()
        rep = mg.Solve(U, F, H.CGInfo(tol, 1e-30, 100), H.NLInfo(1e-10, 1e-30, 50), norm_kind=1)
        assert H.N.NL_STATUS[rep.status] == "converged"
        u = U.copy_back()
        r = f - A @ u
        r[lev.dirichlet] = 0.0
        assert np.linalg.norm(r) <= 1e-10 * np.linalg.norm(f) * 1.0001
        out[tag] = u
        mg.free()
    assert rel(out["mixed"], out["fp64"]) <= 1e-8
    sol = []
    for adjoint, rhs in ((False, f), (True, g)):
        mg = H.CudaMultiLevelSolver(ctx, h, adjoint=adjoint, **common)
        X = mg.new_vector()
        rep = mg.Solve(X, H.CudaVectorInterface.from_global(ctx, rhs),
                       H.CGInfo(1e-12, 1e-30, 200), H.NLInfo(1e-12, 1e-30, 50), norm_kind=1)
        assert H.N.NL_STATUS[rep.status] == "converged"
        sol.append(X.copy_back())
        mg.free()
    u, z = sol
    assert abs(g @ u - z @ f) <= 1e-10 * abs(g @ u)
