"""The colour order for the ILU smoother (ghs_colour_order).

Checks by brute force on the CPU: the order is a permutation; nodes of one
colour are never coupled, so in this order the lower (and the upper)
triangle's dependency graph has at most `ncolours` levels; a uniform Q1
hexahedral level needs 8 colours (the 2x2x2 parity classes), the adaptive
Q1 and Q2 levels of the reference's fixtures more.  The oracle's ILU in the
colour order still equals the reference's algorithm applied to the permuted
matrix (ILU in any order is the identity-order ILU of P A P^T)."""
import numpy as np
import pytest

import oracle_ctypes as O
from gascoignelib_amd.structure import colour_order, uniform_hierarchy


def classes(rowptr, col, perm, nc):
    """colour of each node recovered from the order: a new class starts where
    a node is coupled to an earlier node of the current class"""
    n = len(rowptr) - 1
    colour = -np.ones(n, np.int64)
    c, members = 0, set()
    for v in perm:
        nb = set(col[rowptr[v]:rowptr[v + 1]].tolist()) - {v}
        if nb & members:
            c += 1
            members = set()
        colour[v] = c
        members.add(v)
    return colour


def levels(rowptr, col, perm):
    n = len(rowptr) - 1
    q = np.empty(n, np.int64)
    q[perm] = np.arange(n)
    lf = np.zeros(n, np.int64)
    for i in range(n):
        v = perm[i]
        c = q[col[rowptr[v]:rowptr[v + 1]]]
        lo = c[c < i]
        if len(lo):
            lf[i] = lf[lo].max() + 1
    return int(lf.max()) + 1


def test_uniform_levels_use_eight_colours():
    h = uniform_hierarchy("cube.inp", 3, 4, drop_coarse=1)
    for l in range(h.nlevels):
        lev = h.level(l)
        perm, nc = colour_order(lev.rowptr, lev.col)
        assert np.array_equal(np.sort(perm), np.arange(lev.n))
        assert nc == 8
        colour = classes(lev.rowptr, lev.col, perm, nc)
        assert colour.max() + 1 <= nc
        assert levels(lev.rowptr, lev.col, perm) == nc


@pytest.mark.parametrize("name", ["q1_3d_8_ada", "q2_3d_8_ada", "q1_2d_128"])
def test_fixture_levels(golden, name):
    z = golden(name)
    for l in range(int(z["nlevels"])):
        rp, col = z[f"L{l}_rowptr"], z[f"L{l}_col"]
        n = len(rp) - 1
        if n > 20000:
            continue
        perm, nc = colour_order(rp, col)
        assert np.array_equal(np.sort(perm), np.arange(n))
        colour = classes(rp, col, perm, nc)
        # no coupling inside a colour class
        for v in range(n):
            nb = col[rp[v]:rp[v + 1]]
            nb = nb[nb != v]
            assert not np.any(colour[nb] == colour[v])
        assert levels(rp, col, perm) <= nc


def test_ilu_in_colour_order_is_ilu_of_permuted_matrix(golden):
    """SparseBlockIlu in order p = the identity-order ILU of A(p, p)"""
    z = golden("q1_3d_8_ada")
    rp, col, val = z["L1_rowptr"], z["L1_col"], z["L1_val"]
    n = len(rp) - 1
    perm, _ = colour_order(rp, col)
    q = np.empty(n, np.int64)
    q[perm] = np.arange(n)
    # A(p, p) in CSR with the permuted rows and renumbered columns
    prp, pcol, pval = [0], [], []
    for i in range(n):
        v = perm[i]
        cc = q[col[rp[v]:rp[v + 1]]]
        o = np.argsort(cc, kind="stable")
        pcol += cc[o].tolist()
        pval += val[rp[v]:rp[v + 1]][o].tolist()
        prp.append(len(pcol))
    I1 = O.Ilu(O.Matrix(rp, col, val), perm)
    I2 = O.Ilu(O.Matrix(prp, pcol, pval), None)
    assert np.array_equal(I1.val, I2.val)
    x = np.random.default_rng(0).uniform(-1, 1, n)
    x1, x2 = x.copy(), x[perm].copy()
    I1.solve(x1)
    I2.solve(x2)
    assert np.array_equal(x1[perm], x2)
