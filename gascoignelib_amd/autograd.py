"""Gradients through the HIP forward and adjoint solves: per-cell coefficient
gradients of a diffusion functional, and the gradient of the Stokes channel's
outflow flux with respect to its inflow data (StokesFluxProblem, InflowFlux).

Problem: -div(k grad u) = f on a Q1 box mesh with one coefficient k_c per cell,
homogeneous Dirichlet data, and the functional J(k) = g . u(k) (g = dJ/du, for
J = int u w dx the lumped-mass weights g_i = m_i w(x_i), zero on Dirichlet
nodes).  With A(k) = sum_c k_c K^c (Dirichlet rows and columns eliminated):

    forward:   A u = f                       (MG-preconditioned GMRES, device)
    adjoint:   A^T z = g                     (the solver on the transposed levels)
    gradient:  dJ/dk_c = -z^T K^c u          (gh_dJdk_cells, device)

`CoefficientFunctional` wraps this as a torch.autograd.Function so that the
gradient can be checked with torch's own machinery against central finite
differences and an FP64 CPU adjoint.  It is the gradient-check harness, not a
training path: k lives on the host, every forward assembles on the host.
"""
import numpy as np

from . import hip as H
from .problem import lattice_rhs, lumped_mass, smooth_weight
from .stokes import StokesFluxProblem  # noqa: F401  (re-exported for the torch wrapper)
from .structure import Mesh, assemble_hierarchy


class DiffusionProblem:
    """One mesh hierarchy, right-hand side and functional; solves for any k."""

    def __init__(self, ctx, inp="cube3.inp", dim=3, prerefine=2, drop_coarse=1, seed=2024,
                 wseed=77, prec="f64", krylov="gmres", tol=1e-13, ndirect=400,
                 order="lexicographic"):
        self.ctx = ctx
        self.mesh = Mesh(inp, dim, prerefine, drop_coarse)
        self.dim = dim
        fl = self.mesh.nlevels - 1
        self.coords = self.mesh.coords(fl)
        self.cells = self.mesh.cells(fl)
        self.dirichlet = self.mesh.dirichlet(fl)
        self.f = lattice_rhs(self.coords, seed)
        self.f[self.dirichlet] = 0.0
        self.g = lumped_mass(self.coords, self.cells) * smooth_weight(self.coords, wseed)
        self.g[self.dirichlet] = 0.0
        self.settings = dict(mgtype="V", iterpre=2, iterpost=2, iterexact=10, omega=0.8,
                             mgomega=1.0, ndirect=ndirect, prec=prec, krylov=krylov,
                             order=order)
        self.tol = tol
        self.cellmap = H.CoefficientCells(ctx, dim, self.coords, self.cells)
        self.F = H.CudaVectorInterface.from_global(ctx, self.f)
        self.G = H.CudaVectorInterface.from_global(ctx, self.g)

    @property
    def ncells(self):
        return len(self.cells)

    def hierarchy(self, k):
        return assemble_hierarchy(self.mesh, np.asarray(k, dtype=np.float64))

    def _solve(self, hier, rhs, adjoint):
        mg = H.CudaMultiLevelSolver(self.ctx, hier, adjoint=adjoint, **self.settings)
        X = mg.new_vector()
        lin = self.tol * (0.5 if self.settings["krylov"] == "fgmres" else 1.0)
        rep = mg.Solve(X, rhs, H.CGInfo(lin, 1e-30, 200), H.NLInfo(self.tol, 1e-30, 50),
                       norm_kind=1)
        if H.N.NL_STATUS[rep.status] != "converged":
            raise RuntimeError(f"{'adjoint' if adjoint else 'forward'} solve: "
                               f"{H.N.NL_STATUS[rep.status]}")
        return X, rep

    def state(self, k):
        """(hierarchy, u on the device, report)"""
        hier = self.hierarchy(k)
        U, rep = self._solve(hier, self.F, adjoint=False)
        return hier, U, rep

    def J(self, k):
        _, U, _ = self.state(k)
        return float(self.g @ U.copy_back())

    def gradient(self, hier, U):
        """(z on the device, dJ/dk as numpy)"""
        Z, _ = self._solve(hier, self.G, adjoint=True)
        Gk = H.CudaVectorInterface(self.ctx, self.ncells)
        self.cellmap.dJdk(Z, U, Gk)
        return Z, Gk.copy_back()


try:
    import torch

    class CoefficientFunctional(torch.autograd.Function):
        """J = CoefficientFunctional.apply(k, problem) for a float64 tensor k of
        one coefficient per cell; backward runs the adjoint solve."""

        @staticmethod
        def forward(fctx, k, problem):
            kk = k.detach().cpu().numpy().astype(np.float64)
            hier, U, _ = problem.state(kk)
            fctx.problem = problem
            fctx.hier = hier
            fctx.U = U
            return torch.tensor(float(problem.g @ U.copy_back()), dtype=torch.float64)

        @staticmethod
        def backward(fctx, grad_out):
            _, dk = fctx.problem.gradient(fctx.hier, fctx.U)
            return grad_out * torch.from_numpy(dk), None

    class InflowFlux(torch.autograd.Function):
        """J = InflowFlux.apply(g, problem): outflow flux of the Stokes channel
        for a float64 tensor g of inflow values; backward runs the adjoint."""

        @staticmethod
        def forward(fctx, g, problem):
            fctx.problem = problem
            return torch.tensor(problem.J(g.detach().cpu().numpy().astype(np.float64)),
                                dtype=torch.float64)

        @staticmethod
        def backward(fctx, grad_out):
            dg, _, _ = fctx.problem.gradient()
            return grad_out * torch.from_numpy(dg), None
except ImportError:  # torch is only needed for the gradient check
    CoefficientFunctional = None
    InflowFlux = None
