"""Stokes-LPS block assembly of the structure producer (CPU).

The producer forms the patch element of NavierStokesLps<3> (U = 0) as tensor
products of exact 1D integrals.  It is pinned here against an independent
route: the reference's own recipe evaluated by quadrature - the Q1-patch basis
of BaseQ13dPatch::point (src/Discretization/Q2/baseq13dpatch.h:104-140) at the
2x2x2 Gauss points of every sub-cell, the fluctuation lists of
ElementLpsIntegrator::Projection (src/DofHandler/elementlpsintegrator.h:77-103),
the entries of NavierStokesLps::Matrix / StabMatrix
(src/Problems/navierstokes.h:144-168, 251-268) and lpspoint (:221-227) - on a
non-cubic box so that a swapped direction, a transposed block or a wrong sign
fails.  Further closed forms: the patch structure (125 couplings for interior
corner nodes), the divergence of linear fields, the LPS kernel (trilinear
pressures on a patch), Dirichlet rows and columns, the lifted residual and the
boundary masses.
"""

import numpy as np
import pytest

from gascoignelib_amd.structure import Mesh, stokes_hierarchy


def box_inp(path, lx, ly, lz):
    X = [(0, 0, 0), (lx, 0, 0), (lx, ly, 0), (0, ly, 0),
         (0, 0, lz), (lx, 0, lz), (lx, ly, lz), (0, ly, lz)]
    lines = ["8 7 0 0 0"] + [f"{i} {x!r} {y!r} {z!r}" for i, (x, y, z) in enumerate(X)]
    lines += ["0 0 hex 0 1 2 3 4 5 6 7", "0 2 quad 0 1 2 3", "1 2 quad 5 4 7 6",
              "2 3 quad 1 5 6 2", "3 2 quad 3 2 6 7", "4 1 quad 0 3 7 4", "5 2 quad 1 0 4 5"]
    path.write_text("\n".join(lines) + "\n")
    return str(path)


def dense_blocks(rowptr, col, val, nodes):
    """dense (27*4)^2 matrix of the block entries among `nodes` (local order)"""
    loc = {int(v): k for k, v in enumerate(nodes)}
    m = len(nodes)
    D = np.zeros((4 * m, 4 * m))
    V = val.reshape(-1, 4, 4)
    for i in nodes:
        for q in range(rowptr[i], rowptr[i + 1]):
            j = int(col[q])
            if j in loc:
                a, b = loc[int(i)], loc[j]
                D[4 * a:4 * a + 4, 4 * b:4 * b + 4] = V[q]
    return D


# ElementLpsIntegrator::Projection (elementlpsintegrator.h:77-103): corner ->
# (edge midpoints, face centres) whose test functions enter its fluctuation;
# every corner also takes -1/8 of the centre node 13
PROJ = {0: ((1, 3, 9), (4, 10, 12)), 2: ((1, 5, 11), (4, 10, 14)),
        6: ((3, 7, 15), (4, 16, 12)), 8: ((5, 7, 17), (4, 16, 14)),
        18: ((19, 21, 9), (22, 10, 12)), 20: ((19, 23, 11), (22, 10, 14)),
        24: ((21, 25, 15), (22, 16, 12)), 26: ((23, 25, 17), (22, 16, 14))}


def quadrature_patch(hx, hy, hz, visc, alpha):
    """(27*4)^2 patch element by the reference recipe, evaluated at Gauss points"""
    g = (0.5 - 0.5 / np.sqrt(3.0), 0.5 + 0.5 / np.sqrt(3.0))
    L = np.array([2 * hx, 2 * hy, 2 * hz])
    J = np.prod(L)
    h = J ** (1.0 / 3.0)
    a_lps = alpha / (visc / h / h)
    E = np.zeros((27, 27, 4, 4))

    def hat(k, s):  # 1D Q1-patch function of node k at s in [0,1], and d/ds
        if k == 0:
            return (max(0.0, 1 - 2 * s), -2.0 if s < 0.5 else 0.0)
        if k == 2:
            return (max(0.0, 2 * s - 1), 2.0 if s > 0.5 else 0.0)
        return ((2 * s if s <= 0.5 else 2 - 2 * s), 2.0 if s < 0.5 else -2.0)

    for cz in range(2):
        for cy in range(2):
            for cx in range(2):
                for gz in g:
                    for gy in g:
                        for gx in g:
                            s = (0.5 * (cx + gx), 0.5 * (cy + gy), 0.5 * (cz + gz))
                            w = 0.125 * 0.125 * J
                            N = np.zeros(27)
                            dN = np.zeros((27, 3))
                            for a in range(27):
                                idx = (a % 3, (a // 3) % 3, a // 9)
                                v = [hat(idx[d], s[d]) for d in range(3)]
                                N[a] = v[0][0] * v[1][0] * v[2][0]
                                dN[a, 0] = v[0][1] * v[1][0] * v[2][0] / L[0]
                                dN[a, 1] = v[0][0] * v[1][1] * v[2][0] / L[1]
                                dN[a, 2] = v[0][0] * v[1][0] * v[2][1] / L[2]
                            dP = dN.copy()
                            for c, (edges, faces) in PROJ.items():
                                dP[c] = -0.125 * dN[13]
                                for e in edges:
                                    dP[c] -= 0.5 * dN[e]
                                for f in faces:
                                    dP[c] -= 0.25 * dN[f]
                            for i in range(27):        # test (row)
                                for j in range(27):    # trial (column)
                                    for d in range(3):
                                        E[i, j, 0, d + 1] += w * dN[j, d] * N[i]
                                        E[i, j, d + 1, d + 1] += w * visc * dN[j] @ dN[i]
                                        E[i, j, d + 1, 0] -= w * N[j] * dN[i, d]
                                    E[i, j, 0, 0] += a_lps * (w * dP[j]) @ dP[i]
    return E.transpose(0, 2, 1, 3).reshape(108, 108)


@pytest.fixture(scope="module")
def box(tmp_path_factory):
    return box_inp(tmp_path_factory.mktemp("stokes") / "box.inp", 2.0, 1.0, 0.5)


def test_patch_element_matches_reference_quadrature(box):
    visc, alpha = 0.7, 0.3
    M = Mesh(box, 3, 1, 1)  # one patch of 2x2x2 cells
    rp, col = M.patch_structure(0)
    P = M.patches(0)
    assert P.shape == (1, 27) and sorted(P[0]) == list(range(27))
    assert rp[-1] == 27 * 27
    val, _ = M.assemble_stokes(0, len(col), visc, alpha, (), ())
    got = dense_blocks(rp, col, val, P[0])
    want = quadrature_patch(1.0, 0.5, 0.25, visc, alpha)
    assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()


def test_patches_are_lexicographic_boxes():
    M = Mesh("channel.inp", 3, 3, 1)
    for l in range(M.nlevels):
        X = M.coords(l)
        P = M.patches(l)
        n1 = 2 ** (l + 1)
        assert P.shape == ((n1 // 2) ** 3, 27)
        x0 = X[P[:, 0]]
        hP = X[P[:, 26]] - x0
        assert np.allclose(hP, 2.0 / n1)
        for a in range(27):
            idx = np.array([a % 3, (a // 3) % 3, a // 9])
            assert np.allclose(X[P[:, a]], x0 + idx * hP / 2)
        # every cell belongs to exactly one patch: patches tile the cube
        assert len(np.unique(P[:, 13])) == len(P)


def test_patch_structure_sizes():
    M = Mesh("channel.inp", 3, 3, 1)
    for l in range(M.nlevels):
        p = 2 ** l  # patches per direction
        n, _, _ = M.sizes(l)
        rp, col = M.patch_structure(l)
        assert n == (2 * p + 1) ** 3
        assert rp[-1] == (8 * p + 1) ** 3
        assert np.diff(rp).max() == (125 if p > 1 else 27)
        for i in range(0, n, 37):
            assert np.all(np.diff(col[rp[i]:rp[i + 1]]) > 0)


def test_operator_closed_forms(box):
    visc, alpha = 1.3, 0.1
    M = Mesh(box, 3, 2, 1)
    l = M.nlevels - 1
    rp, col = M.patch_structure(l)
    val, _ = M.assemble_stokes(l, len(col), visc, alpha, (), ())
    n = len(rp) - 1
    X = M.coords(l)
    rows = np.repeat(np.arange(n), np.diff(rp))
    V = val.reshape(-1, 4, 4)

    def apply(u):  # (n, 4) -> (n, 4)
        out = np.zeros((n, 4))
        np.add.at(out, rows, np.einsum("kst,kt->ks", V, u[col]))
        return out

    # linear velocity u = (x, 2y, -z): int div(u) phi_i = (1 + 2 - 1) int phi_i
    u = np.zeros((n, 4))
    u[:, 1], u[:, 2], u[:, 3] = X[:, 0], 2 * X[:, 1], -X[:, 2]
    div = apply(u)[:, 0]
    # int phi_i from the divergence of the field (x, 0, 0)
    e = np.zeros((n, 4))
    e[:, 1] = X[:, 0]
    mass = apply(e)[:, 0]
    assert abs(mass.sum() - 1.0) < 1e-13  # box volume 2 * 1 * 0.5
    assert np.allclose(div, 2.0 * mass, atol=1e-14)
    # constant velocity: no divergence, no viscous reaction
    c = np.zeros((n, 4))
    c[:, 1:] = (1.0, -2.0, 0.5)
    assert np.abs(apply(c)).max() < 1e-13
    # constant pressure: no LPS reaction, and -int dphi_i/dx_d vanishes inside
    c = np.zeros((n, 4))
    c[:, 0] = 1.0
    r = apply(c)
    inner = np.all((X > 1e-12) & (X < np.array([2.0, 1.0, 0.5]) - 1e-12), axis=1)
    assert np.abs(r[:, 0]).max() < 1e-13 and np.abs(r[inner]).max() < 1e-13
    # the LPS term vanishes on pressures that are trilinear on every patch
    p = np.zeros((n, 4))
    p[:, 0] = X[:, 0] * X[:, 1] * X[:, 2] + 0.5 * X[:, 1]
    r = apply(p)
    assert np.abs(r[:, 0]).max() < 1e-13
    # the pressure-pressure block is symmetric positive semidefinite
    Lpp = V[:, 0, 0]
    T = {(int(i), int(j)): v for i, j, v in zip(rows, col, Lpp)}
    assert all(abs(v - T[(j, i)]) <= 1e-15 * (1 + abs(v)) for (i, j), v in T.items())
    # the velocity blocks are visc * the Q1 Laplacian (rows sum to zero)
    s = np.zeros(n)
    np.add.at(s, rows, V[:, 1, 1])
    assert np.abs(s).max() < 1e-12
    assert np.array_equal(V[:, 1, 1], V[:, 2, 2]) and np.array_equal(V[:, 1, 1], V[:, 3, 3])
    # gradient and divergence are negative transposes of each other
    for d in range(3):
        G = {(int(i), int(j)): v for i, j, v in zip(rows, col, V[:, d + 1, 0])}
        Dv = {(int(i), int(j)): v for i, j, v in zip(rows, col, V[:, 0, d + 1])}
        assert all(abs(v + Dv[(j, i)]) < 1e-15 for (i, j), v in G.items())


def test_dirichlet_and_lifted_residual(box):
    M = Mesh(box, 3, 2, 1)
    l = M.nlevels - 1
    rp, col = M.patch_structure(l)
    n = len(rp) - 1
    full, _ = M.assemble_stokes(l, len(col), 1.0, 0.1, (), ())
    rng = np.random.default_rng(3)
    u0 = rng.uniform(-1, 1, (n, 4))
    val, rhs = M.assemble_stokes(l, len(col), 1.0, 0.1, (1, 2), (1, 2, 3), u0)
    dn = np.zeros(n, bool)
    dn[M.dirichlet(l, (1, 2))] = True
    rows = np.repeat(np.arange(n), np.diff(rp))
    F, V = full.reshape(-1, 4, 4), val.reshape(-1, 4, 4)
    Au = np.zeros((n, 4))
    np.add.at(Au, rows, np.einsum("kst,kt->ks", F, u0[col]))
    want = -Au
    want[dn, 1:] = 0.0
    assert np.allclose(rhs.reshape(n, 4), want, rtol=0, atol=1e-13)
    for k in range(len(col)):
        i, j = rows[k], col[k]
        ref = F[k].copy()
        if dn[i]:
            ref[1:, :] = 0.0
        if dn[j]:
            ref[:, 1:] = 0.0
        if dn[i] and i == j:
            ref[1, 1] = ref[2, 2] = ref[3, 3] = 1.0
        assert np.array_equal(V[k], ref)


def test_face_masses_and_hierarchy():
    h = stokes_hierarchy(3, seed=11)
    assert 0.5 <= h.amplitude < 1.5
    assert abs(h.mass_in.sum() - 1.0) < 1e-14 and abs(h.mass_out.sum() - 1.0) < 1e-14
    X = h.fine.coords
    assert np.all(h.mass_in[X[:, 0] > 0] == 0) and np.all(h.mass_out[X[:, 0] < 1] == 0)
    # the lifting carries the parabolic profile on the inflow face only
    inflow = h.mesh.dirichlet(h.nlevels - 1, (1,))
    y, z = X[inflow, 1], X[inflow, 2]
    assert np.allclose(h.u0[inflow, 1], h.amplitude * 16 * y * (1 - y) * z * (1 - z))
    assert np.count_nonzero(h.u0) == np.count_nonzero(h.u0[inflow, 1])
    for lev, P in zip(h.levels, h.patches):
        assert lev.ncomp == 4 and list(lev.dirichlet_comps) == [1, 2, 3]
        assert len(lev.val) == 16 * lev.nnz and P.shape[1] == 27


def test_operator_equals_compiled_reference(golden):
    """The whole Stokes-LPS operator of every level against the compiled
    reference's own assembly (fixture stokes_3d_8: NavierStokesLps<3> at U = 0
    on CGDiscQ13dLps, visc 1, alpha 0.1, velocity Dirichlet rows of colours 1
    and 2, channel 8^3): under the permutation that matches node coordinates,
    the block pattern is identical and every entry agrees to 1e-13 of the
    largest (the two assemble with different quadrature and summation order)."""
    import scipy.sparse as sp
    z = golden("stokes_3d_8")
    nl = int(z["nlevels"])
    h = stokes_hierarchy(3, drop_coarse=1)
    checked = 0
    for m in range(nl):
        p = f"L{m}_"
        if p + "stokes_val" not in z.files:
            continue
        xr = z[p + "coords"]
        lev = next(h.level(s) for s in range(h.nlevels) if h.level(s).n == len(xr))
        pos = {tuple(r): i for i, r in enumerate(np.round(lev.coords, 10))}
        perm = np.array([pos[tuple(r)] for r in np.round(xr, 10)])
        assert len(np.unique(perm)) == len(perm)
        rp, col, val = z[p + "stokes_rowptr"], z[p + "stokes_col"], z[p + "stokes_val"]
        n = len(rp) - 1

        def scalar(rp, col, val, n):
            V = val.reshape(-1, 4, 4)
            rows = np.repeat(np.arange(n), np.diff(rp))
            R = (4 * rows[:, None, None] + np.arange(4)[None, :, None]).repeat(4, axis=2)
            Cc = (4 * col[:, None, None] + np.arange(4)[None, None, :]).repeat(4, axis=1)
            return sp.csr_matrix((V.ravel(), (R.ravel(), Cc.ravel())), shape=(4 * n, 4 * n))

        Ar = scalar(rp, col, val, n)
        Am = scalar(lev.rowptr, lev.col, lev.val, lev.n)
        P4 = (4 * perm[:, None] + np.arange(4)).ravel()
        assert abs(Am[P4][:, P4] - Ar).max() <= 1e-13 * abs(Ar).max()
        Pr = sp.csr_matrix((np.ones(len(col)), col, rp))
        Pm = sp.csr_matrix((np.ones(len(lev.col)), lev.col, lev.rowptr))[perm][:, perm]
        assert abs(Pr - Pm).max() == 0
        # velocity Dirichlet nodes: colours 1 and 2
        bnd = np.union1d(z[p + "bnd1"], z[p + "bnd2"])
        assert np.array_equal(np.sort(perm[bnd]), np.sort(lev.dirichlet))
        checked += 1
    assert checked == nl
