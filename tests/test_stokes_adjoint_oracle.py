"""Outflow-flux gradient of the Stokes-LPS channel with respect to the inflow
data, on the CPU with the oracle (the restatement the GPU gradient is checked
against).

J(g) = j . u, u = u0(g) + d, A_D d = -P_F A u0(g); adjoint A_D^T z = P_F j,
dJ/du0 = j - A^T z.  Pinned by what the discretisation fixes, not by the
formula: discrete mass conservation makes J(g) = int_in u0 . e_x ds, so
dJ/dg_i is the inflow-face mass of node i, and dJ/du0 vanishes on every
unconstrained entry (there it reads j_F - A_FF^T z_F, the adjoint equation);
J is linear in g, so central differences of oracle forward solves match the
gradient up to the solver tolerance.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle_ctypes as O
from gascoignelib_amd.problem import Hierarchy, Level
from gascoignelib_amd.structure import stokes_hierarchy

MGP = dict(mgtype="V", iterpre=2, iterpost=2, iterexact=4, omega=0.8, mgomega=1.0, ndirect=200)


def bsr(rowptr, col, val, n):
    return sp.bsr_matrix((np.asarray(val).reshape(-1, 4, 4), col, rowptr), shape=(4 * n, 4 * n))


def transposed(h):
    levels = []
    for lev in h.levels:
        T = bsr(lev.rowptr, lev.col, lev.val, lev.n).T.tobsr(blocksize=(4, 4))
        T.sort_indices()
        levels.append(Level(lev.n, 4, T.indptr.astype(np.int32), T.indices.astype(np.int32),
                            np.ascontiguousarray(T.data.reshape(-1)), lev.transfer, lev.hn_kind,
                            dirichlet=lev.dirichlet, dirichlet_comps=lev.dirichlet_comps,
                            coords=lev.coords))
    return Hierarchy(levels, h.dim, h.name + ":T")


@pytest.fixture(scope="module")
def setup():
    h = stokes_hierarchy(3, seed=5, with_full=True)
    fine = h.fine
    n = fine.n
    A = bsr(fine.rowptr, fine.col, h.full_val, n)
    constrained = np.zeros((n, 4), bool)
    constrained[fine.dirichlet, 1:] = True
    constrained = constrained.reshape(-1)
    j = np.zeros((n, 4))
    j[:, 1] = h.mass_out
    j = j.reshape(-1)
    return h, A, constrained, j


def forward(h, A, constrained, u0):
    r = -(A @ u0)
    r[constrained] = 0.0
    om = O.MG(h, smoother="vanka", patches=h.patches, **MGP)
    d = np.zeros_like(r)
    rc, _ = om.newton(d, r, O.cginfo(1e-13, 1e-30, 200), O.nlinfo(1e-13, 1e-30, 50),
                      norm_kind=1)
    assert rc == 0
    return u0 + d


def test_lifting_residual_matches_assembly(setup):
    h, A, constrained, _ = setup
    r = -(A @ h.u0.reshape(-1))
    r[constrained] = 0.0
    assert np.abs(r - h.rhs).max() <= 1e-13 * np.abs(r).max()


def test_flux_gradient_closed_form(setup):
    h, A, constrained, j = setup
    jF = j.copy()
    jF[constrained] = 0.0
    om = O.MG(transposed(h), smoother="vanka", patches=h.patches, **MGP)
    z = np.zeros_like(jF)
    rc, _ = om.newton(z, jF, O.cginfo(1e-13, 1e-30, 200), O.nlinfo(1e-13, 1e-30, 50),
                      norm_kind=1)
    assert rc == 0
    w = j - A.T @ z
    grad = w.reshape(-1, 4)[h.inflow_nodes, 1]
    want = h.mass_in[h.inflow_nodes]
    assert np.abs(grad - want).max() <= 1e-9 * want.max()
    assert np.abs(w[~constrained]).max() <= 1e-9 * np.abs(j).max()
    # the forward flux: outflow equals inflow
    u = forward(h, A, constrained, h.u0.reshape(-1))
    fin = h.mass_in @ h.u0[:, 1]
    assert abs(j @ u - fin) <= 1e-9 * fin


def test_flux_finite_differences(setup):
    h, A, constrained, j = setup
    g = h.u0[h.inflow_nodes, 1].copy()
    rng = np.random.default_rng(3)
    pick = rng.choice(len(g), 4, replace=False)
    eps = 1e-3
    for k in pick:
        vals = []
        for s in (1.0, -1.0):
            u0 = h.u0.copy()
            u0[h.inflow_nodes[k], 1] = g[k] + s * eps
            vals.append(j @ forward(h, A, constrained, u0.reshape(-1)))
        fd = (vals[0] - vals[1]) / (2 * eps)
        assert abs(fd - h.mass_in[h.inflow_nodes[k]]) <= 1e-7 * h.mass_in.max()
