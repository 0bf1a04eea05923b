"""f3r_vec_dot_noalloc, the entry point that allocates nothing (its
reduction scratch is a static device array): through the C ABI, against
f3r_vec_dot on the same vectors, bit for bit, on two calls in a row (the
second with a smaller input), for every precision pair it accepts, and its
argument checks."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _vec(F, ctx, arr, dt):
    return F.vector.from_numpy(ctx, arr, dt)


def _dot(lib, a, b):
    d = ctypes.c_double()
    assert lib.f3r_vec_dot(a.handle, b.handle, ctypes.byref(d)) == 0
    return d.value


@pytest.mark.parametrize("da,db", [("f64", "f64"), ("f32", "f32"), ("f16", "f16"), ("f64", "f16"),
                                   ("f32", "f16")])
def test_dot_noalloc_equals_dot(da, db):
    import f3r_amd as F
    from f3r_amd._capi import lib
    ctx = F.Context(0)
    out = F.vector(ctx, "f64", 3).fill(-7.0)
    rng = np.random.default_rng(4)
    for n in (3_000_017, 1_001):   # the second call smaller than the first
        print(f"dot_noalloc {da}x{db} n={n}", flush=True)
        x = rng.uniform(-1, 1, n)
        y = rng.uniform(-1, 1, n)
        a, b = _vec(F, ctx, x, da), _vec(F, ctx, y, db)
        assert lib.f3r_vec_dot_noalloc(a.handle, b.handle, out.handle) == 0
        got = out.download()
        ref = _dot(lib, a, b)
        assert got[0] == ref, (n, got[0], ref)
        assert got[1] == -7.0 and got[2] == -7.0   # only out[0] is written
        ctx.sync()
    ctx.close()


def test_dot_noalloc_argument_checks():
    import f3r_amd as F
    from f3r_amd._capi import lib
    ctx = F.Context(0)
    a = F.vector(ctx, "f64", 10).fill(1.0)
    b = F.vector(ctx, "f64", 11).fill(1.0)
    o64 = F.vector(ctx, "f64", 1)
    o32 = F.vector(ctx, "f32", 1)
    i32 = F.vector(ctx, "i32", 10)
    assert lib.f3r_vec_dot_noalloc(a.handle, b.handle, o64.handle) == -1      # lengths differ
    assert lib.f3r_vec_dot_noalloc(a.handle, a.handle, o32.handle) == -1      # out not fp64
    assert lib.f3r_vec_dot_noalloc(a.handle, i32.handle, o64.handle) == -4    # int32 data
    assert lib.f3r_vec_dot_noalloc(None, a.handle, o64.handle) == -1
    # after the refused calls the scratch still gives the right sum
    assert lib.f3r_vec_dot_noalloc(a.handle, a.handle, o64.handle) == 0
    assert o64.download()[0] == 10.0
    ctx.close()
