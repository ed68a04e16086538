"""Structure producer (libgh_structure.so) against the compiled reference.

* CSR structure built from Gascoigne's own cells equals the reference's
  ColumnStencil bit for bit.
* Q1 assembly on Gascoigne's own mesh equals CGDisc::Matrix + dirichlet to
  rounding.
* The producer's own refinement gives the reference's hierarchy up to the
  numbering: same level sizes, and under the coordinate-matching permutation
  the same CSR pattern, matrix, transfer operators and Dirichlet sets.
"""
import numpy as np
import pytest
import scipy.sparse as sp

from gascoignelib_amd import native as N
from gascoignelib_amd.structure import uniform_hierarchy

UNIFORM = [("q1_2d_128", "square.inp", 2, 7), ("q1_3d_16", "cube.inp", 3, 4),
           ("q1_3d_cube3_12", "cube3.inp", 3, 2)]


def structure_from_cells(n, cells):
    cells = np.ascontiguousarray(cells, dtype=np.int32)
    rp = np.empty(n + 1, np.int32)
    N.check_s(N.structure().ghs_structure_from_cells(n, len(cells), cells.shape[1],
                                                     N.ptr(cells, N.i32p), N.ptr(rp, N.i32p), None))
    col = np.empty(rp[-1], np.int32)
    N.check_s(N.structure().ghs_structure_from_cells(n, len(cells), cells.shape[1],
                                                     N.ptr(cells, N.i32p), N.ptr(rp, N.i32p),
                                                     N.ptr(col, N.i32p)))
    return rp, col


@pytest.mark.parametrize("name", [u[0] for u in UNIFORM])
def test_csr_from_reference_cells_bitexact(golden, name):
    z = golden(name)
    for l in range(int(z["nlevels"])):
        p = f"L{l}_"
        n = len(z[p + "coords"])
        rp, col = structure_from_cells(n, z[p + "cells"])
        assert np.array_equal(rp, z[p + "rowptr"])
        assert np.array_equal(col, z[p + "col"])


@pytest.mark.parametrize("name", [u[0] for u in UNIFORM])
def test_assembly_on_reference_mesh(golden, name):
    z = golden(name)
    dim = int(z["dim"])
    for l in range(int(z["nlevels"])):
        p = f"L{l}_"
        xyz = np.zeros((len(z[p + "coords"]), 3))
        xyz[:, :dim] = z[p + "coords"]
        cells = np.ascontiguousarray(z[p + "cells"], dtype=np.int32)
        rp, col = z[p + "rowptr"], z[p + "col"]
        d = np.ascontiguousarray(z[p + "bnd1"], dtype=np.int32)
        val = np.empty(len(col))
        N.check_s(N.structure().ghs_assemble_on(dim, len(xyz), N.ptr(xyz), len(cells),
                                                N.ptr(cells, N.i32p), None,
                                                N.ptr(np.ascontiguousarray(rp), N.i32p),
                                                N.ptr(np.ascontiguousarray(col), N.i32p), len(d),
                                                N.ptr(d, N.i32p), N.ptr(val)))
        ref = z[p + "val"]
        assert np.allclose(val, ref, rtol=1e-13, atol=1e-15 * np.abs(ref).max())


def match_perm(ref_xyz, my_xyz):
    """perm[i_ref] = i_mine, by coordinates"""
    def key(x):
        return [tuple(r) for r in np.round(x, 10)]
    pos = {k: i for i, k in enumerate(key(my_xyz))}
    return np.array([pos[k] for k in key(ref_xyz)])


@pytest.mark.parametrize("name,inp,dim,pre", UNIFORM)
def test_own_refinement_matches_reference_up_to_numbering(golden, name, inp, dim, pre):
    z = golden(name)
    nl = int(z["nlevels"])
    h = uniform_hierarchy(inp, dim, pre, drop_coarse=pre + 1 - nl)
    assert h.nlevels == nl
    perms = []
    for s in range(nl):
        m = nl - 1 - s
        p = f"L{m}_"
        lev = h.level(s)
        perm = match_perm(z[p + "coords"], lev.coords)
        assert len(np.unique(perm)) == lev.n == len(perm)
        perms.append(perm)
        A_ref = sp.csr_matrix((z[p + "val"], z[p + "col"], z[p + "rowptr"]))
        A_my = sp.csr_matrix((lev.val, lev.col, lev.rowptr))
        A_my_in_ref = A_my[perm][:, perm]
        assert abs(A_my_in_ref - A_ref).max() <= 1e-13 * abs(A_ref).max()
        assert np.array_equal(np.sort(perm[z[p + "bnd1"]]), np.sort(lev.dirichlet))
        # the pattern too (explicit zeros included)
        P_ref = sp.csr_matrix((np.ones(len(z[p + "col"])), z[p + "col"], z[p + "rowptr"]))
        P_my = sp.csr_matrix((np.ones(len(lev.col)), lev.col, lev.rowptr))[perm][:, perm]
        assert abs(P_ref - P_my).max() == 0
    # transfer maps: same parents under the permutations
    for s in range(1, nl):
        m = nl - 1 - s
        p = f"L{m}_"
        pf, pc = perms[s], perms[s - 1]
        inv_f = np.empty_like(pf)
        inv_f[pf] = np.arange(len(pf))

        def as_dict(arr, w, fmap, cmap):
            return {int(fmap[r[0]]): tuple(sorted(int(cmap[c]) for c in r[1:w + 1])) for r in arr}

        c2f, zw, vi, ac = h.level(s).transfer
        ident = np.arange(max(len(pf), len(pc)))
        for arr_ref, arr_my, w in ((z[p + "zweier"], zw, 2), (z[p + "vierer"], vi, 4),
                                   (z[p + "achter"], ac, 8)):
            ref = as_dict(arr_ref, w, pf, pc)
            mine = as_dict(arr_my, w, ident, ident)
            assert ref == mine
        # c2f: reference coarse node i sits at fine node c2f_ref[i]
        assert np.array_equal(pf[z[p + "c2f"]], c2f[pc])
