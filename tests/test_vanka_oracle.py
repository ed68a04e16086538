"""The oracle's Vanka smoother pinned without the reference (whose patch LU is
Eigen's PartialPivLU, absent here): one patch covering the whole mesh makes
the smoother an exact solve; on several patches it equals the average of the
independent dense patch solves (LAPACK) over the patches containing a node."""
import numpy as np
import pytest

import oracle_ctypes as O
from gascoignelib_amd.problem import block_operator, patches_from_cells
from gascoignelib_amd.structure import Mesh, assemble_hierarchy


def level(pre):
    M = Mesh("cube.inp", 3, pre, 0)
    fl = M.nlevels - 1
    return assemble_hierarchy(M).fine, M.cells(fl)


def dense(lev, val, B):
    A = np.zeros((lev.n * B, lev.n * B))
    for i in range(lev.n):
        for q in range(lev.rowptr[i], lev.rowptr[i + 1]):
            j = lev.col[q]
            A[i * B:(i + 1) * B, j * B:(j + 1) * B] = val[q * B * B:(q + 1) * B * B].reshape(B, B)
    return A


def operator(lev, B):
    if B == 1:
        return lev.val
    rng = np.random.default_rng(B)
    G = rng.uniform(-1, 1, (B, B))
    return block_operator(lev, G @ G.T + B * np.eye(B), rng.uniform(-0.5, 0.5, (B, B)))


@pytest.mark.parametrize("B", [1, 4])
def test_single_patch_is_exact_solve(B):
    lev, cells = level(1)          # 2x2x2 cells: one patch, all 27 nodes
    patches = patches_from_cells(cells)
    assert patches.shape == (1, 27)
    val = operator(lev, B)
    V = O.Vanka(O.Matrix(lev.rowptr, lev.col, val, B), patches)
    h = np.random.default_rng(1).uniform(-1, 1, lev.n * B)
    x = h.copy()
    V.solve(x)
    ref = np.linalg.solve(dense(lev, val, B), h)
    assert np.abs(x - ref).max() <= 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("B", [1, 4])
def test_average_of_patch_solves(B):
    lev, cells = level(2)          # 4x4x4 cells: 8 overlapping patches
    patches = patches_from_cells(cells)
    assert patches.shape == (8, 27)
    val = operator(lev, B)
    A = dense(lev, val, B)
    h = np.random.default_rng(2).uniform(-1, 1, lev.n * B)
    acc = np.zeros(lev.n * B)
    cnt = np.zeros(lev.n)
    for pn in patches:
        idx = (pn[:, None] * B + np.arange(B)[None]).ravel()
        acc[idx] += np.linalg.solve(A[np.ix_(idx, idx)], h[idx])
        cnt[pn] += 1
    ref = acc / np.repeat(cnt, B)
    x = h.copy()
    O.Vanka(O.Matrix(lev.rowptr, lev.col, val, B), patches).solve(x)
    assert np.abs(x - ref).max() <= 1e-12 * np.abs(ref).max()


def test_pivoting_is_exercised():
    """a dense patch matrix with a zero leading entry (pivoting needed), in
    scalar and 2x2 block form, is solved exactly"""
    rng = np.random.default_rng(3)
    for B in (1, 2):
        m = 27
        rowptr = (np.arange(m + 1) * m).astype(np.int32)
        col = np.tile(np.arange(m, dtype=np.int32), m)
        A = rng.uniform(-1, 1, (m * B, m * B)) + np.eye(m * B)
        A[0, 0] = 0.0
        val = A.reshape(m, B, m, B).transpose(0, 2, 1, 3).reshape(-1)
        h = rng.uniform(-1, 1, m * B)
        x = h.copy()
        O.Vanka(O.Matrix(rowptr, col, val, B), np.arange(m, dtype=np.int32)[None]).solve(x)
        ref = np.linalg.solve(A, h)
        assert np.abs(x - ref).max() <= 1e-11 * np.abs(ref).max(), B
