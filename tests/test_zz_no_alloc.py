"""gh_jacobi_noalloc (h = D^-1 h through a fixed per-GPU scratch array, no
device allocation) against gh_jacobi, through the C ABI.

Two calls in a row on one context, the second on a smaller input, each
compared bit for bit with gh_jacobi on the same matrix and vector (the scratch
keeps the first call's contents, so a stale entry would show in the second).
Scalar FP64 and FP32 matrices and a 3x3 block matrix; h at exactly the
scratch's capacity is accepted, one entry more is GH_EINVAL and leaves h as
it was.
"""
import ctypes as C

import numpy as np
import pytest

from gascoignelib_amd import hip as H
from gascoignelib_amd import native as N
from gascoignelib_amd.problem import block_operator
from gascoignelib_amd.structure import Mesh, assemble_hierarchy

pytestmark = pytest.mark.gpu

SCRATCH_BYTES = 8 << 20


@pytest.fixture(scope="module")
def ctx():
    c = H.Context(0)
    yield c
    c.close()


def both(ctx, A, h, B, prec):
    """(gh_jacobi result, gh_jacobi_noalloc result) for the same h"""
    lib = N.hip()
    V1 = H.CudaVectorInterface.from_global(ctx, h, B, prec)
    V2 = H.CudaVectorInterface.from_global(ctx, h, B, prec)
    assert lib.gh_jacobi(ctx.h, A.h, V1.h) == N.GH_OK
    rc = lib.gh_jacobi_noalloc(ctx.h, A.h, V2.h)
    assert rc == N.GH_OK, lib.gh_last_error().decode()
    out = V1.copy_back(), V2.copy_back()
    V1.free()
    V2.free()
    return out


def q1_levels(pre):
    return assemble_hierarchy(Mesh("cube.inp", 3, pre, 1)).levels


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_noalloc_jacobi_scalar_two_calls(ctx, prec):
    print(f"[jacobi_noalloc scalar {prec}]", flush=True)
    levels = q1_levels(4)  # 16^3 cells, then its 8^3 level
    rng = np.random.default_rng(3)
    for lev in (levels[-1], levels[-2]):
        assert lev.n * (8 if prec == "f64" else 4) <= SCRATCH_BYTES
        A = H.CudaCSRMatrixInterface(ctx, lev.rowptr, lev.col, lev.val, 1, prec)
        h = rng.uniform(-1, 1, lev.n)
        ref, got = both(ctx, A, h, 1, prec)
        assert np.array_equal(ref, got)
        # and it is D^-1 h (FP64: the reciprocal diagonal, to rounding)
        if prec == "f64":
            d = lev.val[np.flatnonzero(lev.col == np.repeat(np.arange(lev.n), np.diff(lev.rowptr)))]
            np.testing.assert_allclose(got * d, h, rtol=1e-15, atol=0)
        A.free()


def test_noalloc_jacobi_block_two_calls(ctx):
    print("[jacobi_noalloc block B=3]", flush=True)
    levels = q1_levels(3)
    rng = np.random.default_rng(5)
    G = rng.uniform(-1, 1, (3, 3))
    Cm = G @ G.T + 3 * np.eye(3)
    S = rng.uniform(-0.3, 0.3, (3, 3))
    for lev in (levels[-1], levels[-2]):
        A = H.CudaCSRMatrixInterface(ctx, lev.rowptr, lev.col, block_operator(lev, Cm, S), 3)
        h = rng.uniform(-1, 1, lev.n * 3)
        ref, got = both(ctx, A, h, 3, "f64")
        assert np.array_equal(ref, got)
        A.free()


def diagonal(ctx, n):
    rowptr = np.arange(n + 1, dtype=np.int32)
    col = np.arange(n, dtype=np.int32)
    val = 2.0 + (np.arange(n) % 7).astype(np.float64)
    return H.CudaCSRMatrixInterface(ctx, rowptr, col, val, 1), val


def test_noalloc_jacobi_capacity_limit(ctx):
    print("[jacobi_noalloc capacity]", flush=True)
    lib = N.hip()
    cap = SCRATCH_BYTES // 8
    # exactly the capacity: accepted, equal to gh_jacobi
    A, d = diagonal(ctx, cap)
    h = np.random.default_rng(7).uniform(-1, 1, cap)
    ref, got = both(ctx, A, h, 1, "f64")
    assert np.array_equal(ref, got)
    A.free()
    # one entry more: refused on the host before any launch, h untouched
    A, d = diagonal(ctx, cap + 1)
    h = np.random.default_rng(8).uniform(-1, 1, cap + 1)
    V = H.CudaVectorInterface.from_global(ctx, h)
    rc = lib.gh_jacobi_noalloc(ctx.h, A.h, V.h)
    assert rc == N.GH_EINVAL
    assert "scratch" in lib.gh_last_error().decode()
    assert np.array_equal(V.copy_back(), h)
    # a call after the refusal still works (smaller input)
    V.free()
    A.free()
    A, d = diagonal(ctx, 1000)
    h = np.random.default_rng(9).uniform(-1, 1, 1000)
    ref, got = both(ctx, A, h, 1, "f64")
    assert np.array_equal(ref, got)
    np.testing.assert_allclose(got * d, h, rtol=1e-15, atol=0)
    A.free()


def test_noalloc_jacobi_rejects_bad_arguments(ctx):
    lib = N.hip()
    A, _ = diagonal(ctx, 10)
    V = H.CudaVectorInterface(ctx, 11)
    assert lib.gh_jacobi_noalloc(ctx.h, A.h, V.h) == N.GH_EINVAL
    assert lib.gh_jacobi_noalloc(ctx.h, A.h, C.c_void_p()) == N.GH_EINVAL
    V.free()
    A.free()
