"""Whole-vector operations of the device NumericVector (lmh_vector_op,
lmh_vector_reduce) against numpy on the same inputs.

Bars: elementwise results bit-identical to numpy except AXPY (the device may
fuse a x + z into one rounding: within 2^-52 relative per entry); reductions
min / max / linfty bit-identical, sums and norms within 1e-13 relative
(a different, fixed summation order); reductions deterministic (two calls
bit-identical); n = 0 gives the identities."""
import ctypes as C

import numpy as np
import pytest

import libmesh_amd as L

pytestmark = pytest.mark.gpu

N = 100_003  # several reduction blocks and a ragged tail


def arrays(ctx, rng, n):
    x = rng.standard_normal(n)
    y = rng.standard_normal(n) + 3.0  # no zeros for the divisions
    z = rng.standard_normal(n) + 3.0
    return x, y, z, L.DeviceArray(ctx, n, data=x), L.DeviceArray(ctx, n, data=y), L.DeviceArray(ctx, n, data=z)


def op(ctx, code, n, a, xd, yd, zd):
    L.check(L.hip().lmh_vector_op(ctx.h, code, n, a, xd.ptr if xd is not None else None,
                                  yd.ptr if yd is not None else None, zd.ptr))


def red(ctx, code, n, xd, yd=None):
    out = C.c_double()
    L.check(L.hip().lmh_vector_reduce(ctx.h, code, n, xd.ptr if xd is not None else None,
                                      yd.ptr if yd is not None else None, C.byref(out)))
    return out.value


@pytest.mark.parametrize("n", [1, 257, N])
def test_elementwise(ctx, n):
    rng = np.random.default_rng(5 + n)
    a = 0.37
    cases = [
        (L.VEC_SET, lambda x, y, z: np.full_like(z, a)),
        (L.VEC_ADD_SCALAR, lambda x, y, z: z + a),
        (L.VEC_SCALE, lambda x, y, z: z * a),
        (L.VEC_MUL, lambda x, y, z: z * x),
        (L.VEC_DIV, lambda x, y, z: z / y),
        (L.VEC_RECIPROCAL, lambda x, y, z: 1.0 / z),
        (L.VEC_ABS, lambda x, y, z: np.abs(x)),
        (L.VEC_PMUL, lambda x, y, z: x * y),
        (L.VEC_PDIV, lambda x, y, z: x / y),
        (L.VEC_COPY, lambda x, y, z: x.copy()),
    ]
    for code, ref in cases:
        x, y, z, xd, yd, zd = arrays(ctx, rng, n)
        print(f"vector op {code} n={n}", flush=True)
        if code == L.VEC_DIV:  # z /= x with x = y here
            op(ctx, code, n, a, yd, None, zd)
        elif code == L.VEC_ABS:  # in place on x
            zd = xd
            op(ctx, code, n, a, None, None, zd)
        else:
            op(ctx, code, n, a, xd, yd, zd)
        assert np.array_equal(zd.download(), ref(x, y, z)), code
    x, y, z, xd, yd, zd = arrays(ctx, rng, n)
    op(ctx, L.VEC_AXPY, n, a, xd, None, zd)
    exact = z + a * x
    assert np.all(np.abs(zd.download() - exact) <= 2.0 ** -52 * (np.abs(exact) + np.abs(a * x) + np.abs(z)))


@pytest.mark.parametrize("n", [1, 257, N])
def test_reductions(ctx, n):
    rng = np.random.default_rng(11 + n)
    x, y, z, xd, yd, zd = arrays(ctx, rng, n)
    assert red(ctx, L.VEC_MIN, n, xd) == x.min()
    assert red(ctx, L.VEC_MAX, n, xd) == x.max()
    assert red(ctx, L.VEC_LINF, n, xd) == np.abs(x).max()
    close = [
        (L.VEC_SUM, x.sum(), None),
        (L.VEC_L1, np.abs(x).sum(), None),
        (L.VEC_L2, np.sqrt((x * x).sum()), None),
        (L.VEC_DOT, (x * y).sum(), yd),
        (L.VEC_L1_DIFF, np.abs(x - y).sum(), yd),
        (L.VEC_L2_DIFF, np.sqrt(((x - y) ** 2).sum()), yd),
    ]
    for code, ref, other in close:
        v = red(ctx, code, n, xd, other)
        assert abs(v - ref) <= 1e-13 * max(abs(ref), np.abs(x).sum()), (code, v, ref)
        assert red(ctx, code, n, xd, other) == v  # deterministic


def test_empty_and_bad_arguments(ctx):
    xd = L.DeviceArray(ctx, 4, data=np.ones(4))
    assert red(ctx, L.VEC_SUM, 0, None) == 0.0
    assert red(ctx, L.VEC_MIN, 0, None) == np.inf
    assert red(ctx, L.VEC_MAX, 0, None) == -np.inf
    op(ctx, L.VEC_SCALE, 0, 2.0, None, None, xd)  # n = 0: nothing written
    assert np.array_equal(xd.download(), np.ones(4))
    assert L.hip().lmh_vector_op(ctx.h, 99, 4, 1.0, None, None, xd.ptr) != 0
    assert L.hip().lmh_vector_op(ctx.h, L.VEC_AXPY, 4, 1.0, None, None, xd.ptr) != 0  # x missing
    out = C.c_double()
    assert L.hip().lmh_vector_reduce(ctx.h, L.VEC_DOT, 4, xd.ptr, None, C.byref(out)) != 0  # y missing
