"""The oracle's ILU(0) (SparseBlockIlu, src/LinAlg/sparseblockilu.xx) against
the compiled reference: on Gascoigne's own meshes (Q1 2D, Q1 3D, the
3x3x3-hex cube, Q1 and Q2 with hanging nodes), in the identity order and in
the reversed order, the permuted structure (rows, ascending columns,
diagonal positions), the factors after compute_ilu and one solve of a seeded
vector are equal bit for bit.  Mathematics on top: for a tridiagonal matrix
ILU(0) is the exact LU, so its solve is A^-1."""
import numpy as np
import pytest

import oracle_ctypes as O

CASES = ["q1_2d_128", "q1_3d_16", "q1_3d_cube3_12", "q1_3d_8_ada", "q2_3d_8_ada"]


@pytest.mark.parametrize("name", CASES)
def test_ilu_equals_reference(golden, name):
    z = golden(name)
    checked = 0
    for l in range(int(z["nlevels"])):
        p = f"L{l}_"
        if p + "ilu0_val" not in z.files:
            continue
        A = O.Matrix(z[p + "rowptr"], z[p + "col"], z[p + "val"])
        n = A.n
        for kind in (0, 1):
            q = f"{p}ilu{kind}_"
            perm = np.arange(n) if kind == 0 else np.arange(n)[::-1]
            I = O.Ilu(A, perm)
            assert np.array_equal(I.rowptr, z[q + "rowptr"])
            assert np.array_equal(I.col, z[q + "col"])
            assert np.array_equal(I.diag, z[q + "diag"])
            assert np.array_equal(I.val, z[q + "val"])
            x = z[p + "ilu_in"].copy()
            I.solve(x)
            assert np.array_equal(x, z[q + "solve"])
            checked += 1
    assert checked >= 2


def test_ilu_of_tridiagonal_is_exact():
    n = 50
    rng = np.random.default_rng(1)
    d = 4.0 + rng.uniform(0, 1, n)
    e = rng.uniform(-1, 1, n - 1)
    rowptr, col, val = [0], [], []
    for i in range(n):
        for j, v in ((i - 1, e[i - 1] if i else 0), (i, d[i]), (i + 1, e[i] if i < n - 1 else 0)):
            if 0 <= j < n:
                col.append(j)
                val.append(v)
        rowptr.append(len(col))
    A = O.Matrix(rowptr, col, val)
    dense = np.zeros((n, n))
    for i in range(n):
        for q in range(rowptr[i], rowptr[i + 1]):
            dense[i, col[q]] = val[q]
    b = rng.uniform(-1, 1, n)
    for perm in (None, np.arange(n)[::-1]):
        x = b.copy()
        O.Ilu(A, perm).solve(x)
        np.testing.assert_allclose(dense @ x, b, rtol=0, atol=1e-13)
