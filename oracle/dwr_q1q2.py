"""Oracle (test infrastructure only: imported by tests/ and nothing else) for
the dual-weighted residual estimator of Q1 solutions with the patch-wise
higher-order interpolation of the adjoint (DWR "Q1Q2").

The reference has no implementation to follow: StdLoop::Estimator
(src/Solver/stdloop.cc:221-273) aborts for every estimator type, and
StdSolver::AdjointForm (src/Solver/stdsolver.cc:1068-1090) aborts too.  What
is restated here is the estimator's mathematics (Becker & Rannacher 2001,
Acta Numerica 10, sections 3-4: J(u) - J(u_h) = rho(u_h)(z - i_h z), with z
approximated by the triquadratic interpolant of the discrete adjoint on the
patches of the refinement), in a form independent of the device kernel's:

  eta_K = int_K f w - int_K k grad u_h . grad w,    w = I2 z_h - z_h,

for every cell K of a patch P (the 2^d children of one coarser cell, 3^d
nodes): u_h and z_h are the Q1 (multilinear) finite-element functions of the
nodal vectors u and z, I2 z_h the Q2 (triquadratic / biquadratic) Lagrange
interpolant of z's values at the patch's 3^d nodes, f the Q2 interpolant of
the nodal values f on the patch, k a constant per cell.  Patches are
axis-aligned boxes whose middle nodes are midpoints.

Here every integral is a Kronecker product of 1D matrices over the two
halves of the patch's reference interval [0, 2] (exact Gauss integration in
1D, with numpy's Legendre nodes); the device kernel instead evaluates the
integrands at 3^d Gauss points per cell.  Parity pins (tests/test_dwr_oracle.py):
eta = 0 for a z whose patch values are multilinear; sum eta = 0 for a linear u
with f = 0 and z = 0 on the boundary; the efficiency index against the exact
J(u) of a smooth problem tends to 1."""
import numpy as np

# 1D functions on the patch's reference interval [0, 2], nodes 0, 1, 2
_Q2 = [lambda t: 0.5 * (t - 1) * (t - 2), lambda t: t * (2 - t), lambda t: 0.5 * t * (t - 1)]
_DQ2 = [lambda t: t - 1.5, lambda t: 2 - 2 * t, lambda t: t - 0.5]
_Q1 = [lambda t: np.clip(1 - t, 0, None), lambda t: 1 - np.abs(t - 1),
       lambda t: np.clip(t - 1, 0, None)]
_DQ1 = [lambda t: np.where(t < 1, -1.0, 0.0), lambda t: np.where(t < 1, 1.0, -1.0),
        lambda t: np.where(t < 1, 0.0, 1.0)]


def _half_matrices(s):
    """1D integrals over the half [s, s+1] of the reference interval:
    M[a][b] = int A_a B_b for the pairs (A, B) in (Q2,Q2), (Q2,Q1), (Q1,Q2),
    (Q1,Q1) and the same for the derivatives"""
    xg, wg = np.polynomial.legendre.leggauss(4)  # exact for the degree-4 products
    t = s + 0.5 * (xg + 1)
    w = 0.5 * wg
    # the Q1 derivatives are constant on a half: take them inside it
    tm = np.full_like(t, s + 0.5)
    val = {"2": np.array([f(t) for f in _Q2]), "1": np.array([f(t) for f in _Q1])}
    der = {"2": np.array([f(t) for f in _DQ2]), "1": np.array([f(tm) for f in _DQ1])}
    M = {a + b: (val[a] * w) @ val[b].T for a in "21" for b in "21"}
    D = {a + b: (der[a] * w) @ der[b].T for a in "21" for b in "21"}
    return M, D


_HALF = [_half_matrices(0), _half_matrices(1)]


def _kron(mats):
    out = np.ones((1, 1))
    for m in mats:  # the first factor is the slowest index: z, y, x order
        out = np.kron(out, m)
    return out


def cell_matrices(dim, h):
    """For each of the 2^dim cells of a patch with cell sizes h (per axis),
    the matrices F, S with  int_K (I2 f) w = f^T F z  and
    int_K grad u_h . grad w = u^T S z  (w = I2 z - z_h), on the patch's 3^dim
    nodes in lexicographic order (x fastest)"""
    out = []
    for cell in range(2 ** dim):
        sub = [(cell >> d) & 1 for d in range(dim)]
        Ms = [_HALF[sub[d]][0] for d in range(dim)]
        Ds = [_HALF[sub[d]][1] for d in range(dim)]
        vol = float(np.prod(h))
        order = list(range(dim))[::-1]  # kron factors: z (slowest) .. x (fastest)
        F = vol * (_kron([Ms[d]["22"] for d in order]) - _kron([Ms[d]["21"] for d in order]))
        S = np.zeros_like(F)
        for g in range(dim):
            fac2 = [Ds[d]["12"] if d == g else Ms[d]["12"] for d in order]
            fac1 = [Ds[d]["11"] if d == g else Ms[d]["11"] for d in order]
            S += vol / h[g] ** 2 * (_kron(fac2) - _kron(fac1))
        out.append((F, S))
    return out


def estimate(dim, coords, patches, u, z, f, k=None):
    """eta per cell, (npatches, 2^dim) in the patches' order and each patch's
    cells in lexicographic order.  coords (nnodes, >= dim); patches
    (npatches, 3^dim) node indices, lexicographic (x fastest)."""
    patches = np.asarray(patches)
    npat = len(patches)
    nc = 2 ** dim
    k = np.ones((npat, nc)) if k is None else np.asarray(k, dtype=np.float64).reshape(npat, nc)
    eta = np.empty((npat, nc))
    cache = {}
    for p in range(npat):
        nodes = patches[p]
        x = coords[nodes][:, :dim]
        h = (x[-1] - x[0]) / 2.0
        key = tuple(np.round(h, 15))
        if key not in cache:
            cache[key] = cell_matrices(dim, h)
        for c, (F, S) in enumerate(cache[key]):
            eta[p, c] = f[nodes] @ F @ z[nodes] - k[p, c] * (u[nodes] @ S @ z[nodes])
    return eta


def load_vector(dim, coords, patches, f, nnodes):
    """b_i = int (I2 f) phi_i, the Q1 load vector of the patch-wise Q2
    interpolant of the nodal values f"""
    b = np.zeros(nnodes)
    cache = {}
    for nodes in np.asarray(patches):
        x = coords[nodes][:, :dim]
        h = (x[-1] - x[0]) / 2.0
        key = tuple(np.round(h, 15))
        if key not in cache:
            order = list(range(dim))[::-1]
            mats = []
            for cell in range(2 ** dim):
                sub = [(cell >> d) & 1 for d in range(dim)]
                mats.append(float(np.prod(h)) * _kron([_HALF[sub[d]][0]["12"] for d in order]))
            cache[key] = sum(mats)  # (Q1 row, Q2 col): int phi_i psi_j over the patch
        b[nodes] += cache[key] @ f[nodes]
    return b
