"""Outflow flux of the Stokes-LPS channel and its gradient with respect to the
inflow data, through the HIP forward and adjoint solves (no torch here; the
torch wrapper is gascoignelib_amd.autograd.InflowFlux)."""
import numpy as np

from . import hip as H


class StokesFluxProblem:
    """Outflow flux of the Stokes-LPS channel as a function of the inflow data.

    g = u_x at the inflow nodes (the others keep the lifting's values); the
    lifting u0(g) carries g, the solution is u = u0 + d with
    A_D d = -P_F A u0 (A the operator without Dirichlet conditions, A_D with
    them, P_F zeroing the constrained entries), and J = j . u with j the
    outflow-face masses on the u_x entries (J = int_out u . n ds).
    Adjoint: A_D^T z = P_F j (the solver on the transposed levels, Vanka
    factors of the transposed patch matrices), then
        dJ/du0 = j - A^T z,     dJ/dg_i = (dJ/du0)(node_i, u_x).
    Discrete mass conservation fixes the answer: the divergence rows of all
    nodes sum to the boundary flux and the LPS term vanishes on constants, so
    J = int_in u0 . e_x ds and dJ/dg_i = int_in phi_i ds."""

    def __init__(self, ctx, hier, prec="f64", krylov="gmres", tol=1e-12, order="lexicographic",
                 omega=0.8, ndirect=200, smooth=(2, 2), vmax=None):
        if hier.full_val is None:
            raise ValueError("StokesFluxProblem needs stokes_hierarchy(..., with_full=True)")
        self.ctx = ctx
        self.h = hier
        fine = hier.fine
        self.n = fine.n
        self.nodes = np.asarray(hier.inflow_nodes, dtype=np.int64)
        common = dict(mgtype="V", iterpre=smooth[0], iterpost=smooth[1], iterexact=4,
                      omega=omega, mgomega=1.0,
                      ndirect=ndirect, linear_smooth="vanka", patches=hier.patches, order=order,
                      prec=prec, krylov=krylov)
        self.mg = H.CudaMultiLevelSolver(ctx, hier, **common)
        self.mgT = H.CudaMultiLevelSolver(ctx, hier, adjoint=True, **common)
        self.A = H.CudaCSRMatrixInterface(ctx, fine.rowptr, fine.col, hier.full_val, 4)
        self.mask = self.mg.fine_mask
        self.tol = tol
        # flexible GMRES: its tolerance is measured from the starting residual
        # and its estimate is the true residual, so a margin of 0.9 suffices;
        # restart length 20 (the reference's GMRES keeps 10)
        self.lin = tol * (0.9 if krylov == "fgmres" else 1.0)
        self.vmax = vmax or (20 if krylov == "fgmres" else 10)
        j = np.zeros((self.n, 4))
        j[:, 1] = hier.mass_out
        self.j = j.reshape(-1)
        self.J_ = H.CudaVectorInterface.from_global(ctx, self.j, 4)
        self.JF = H.CudaVectorInterface.from_global(ctx, self.j, 4)
        self.mask.zero(self.JF)
        self.base = hier.u0.copy()

    def lifting(self, g):
        u0 = self.base.copy()
        u0[self.nodes, 1] = g
        return u0.reshape(-1)

    def _solve(self, mg, rhs):
        X = mg.new_vector()
        rep = mg.Solve(X, rhs, H.CGInfo(self.lin, 1e-30, 200), H.NLInfo(self.tol, 1e-30, 50),
                       vmax=self.vmax, norm_kind=1)
        if H.N.NL_STATUS[rep.status] != "converged":
            raise RuntimeError(f"Stokes solve: {H.N.NL_STATUS[rep.status]}")
        return X, rep

    def state(self, g):
        """(u = u0 + d as numpy, report)"""
        u0 = self.lifting(g)
        U0 = H.CudaVectorInterface.from_global(self.ctx, u0, 4)
        R = H.CudaVectorInterface(self.ctx, self.n, 4)
        self.A.vmult(R, U0, -1.0, 0.0)
        self.mask.zero(R)
        D, rep = self._solve(self.mg, R)
        D.add(1.0, U0)
        return D.copy_back(), rep

    def J(self, g):
        u, _ = self.state(g)
        return float(self.j @ u)

    def gradient(self):
        """dJ/dg (numpy) and the adjoint state z; J is linear in g, so the
        gradient does not depend on g."""
        Z, rep = self._solve(self.mgT, self.JF)
        W = H.CudaVectorInterface.from_global(self.ctx, self.j, 4)
        self.A.vmult_transpose(W, Z, -1.0, 1.0)
        w = W.copy_back().reshape(-1, 4)
        return w[self.nodes, 1].copy(), Z, rep
