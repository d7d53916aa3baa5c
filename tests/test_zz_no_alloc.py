"""fem_vec_dot_noalloc (the allocation-free scalar product) against
fem_vec_dot, through the C ABI: two calls in a row, the second on shorter
vectors, in FP64 and mixed FP32/FP64."""
import ctypes as C

import numpy as np
import pytest

import dunefem_amd as fa

pytestmark = pytest.mark.gpu


def _dot(fn, x, y):
    out = C.c_double()
    rc = fn(x.h, y.h, C.byref(out))
    assert rc == 0, fa._lib.fem_last_error().decode()
    return out.value


def test_noalloc_dot_matches_dot_on_consecutive_calls(ctx):
    L = fa.load_library()
    rng = np.random.default_rng(0)
    for n, px, py in ((3_000_001, fa.F64, fa.F64), (1000, fa.F64, fa.F64), (777_777, fa.F32, fa.F64),
                      (65, fa.F32, fa.F32), (0, fa.F64, fa.F64)):
        a, b = rng.normal(size=n), rng.normal(size=n)
        x = fa.DofVector(ctx, n, px, a if n else None)
        y = fa.DofVector(ctx, n, py, b if n else None)
        ref = _dot(L.fem_vec_dot, x, y)
        got = _dot(L.fem_vec_dot_noalloc, x, y)
        # the same kernel on the same fixed grid: bitwise equal
        assert got == ref, (n, got, ref)
        if n:
            ah = a.astype(np.float32).astype(np.float64) if px == fa.F32 else a
            bh = b.astype(np.float32).astype(np.float64) if py == fa.F32 else b
            assert abs(got - ah @ bh) <= 1e-12 * np.abs(ah) @ np.abs(bh)
        else:
            assert got == 0.0
    # a size mismatch is an argument error, not a device fault
    x, y = fa.DofVector(ctx, 10, fa.F64), fa.DofVector(ctx, 11, fa.F64)
    out = C.c_double()
    assert L.fem_vec_dot_noalloc(x.h, y.h, C.byref(out)) == -1
