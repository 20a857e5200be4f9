"""GPU: the projection's division (a reciprocal step shared by x/w and y/w
where the operands allow it, the compiler's division elsewhere) returns the
IEEE-rounded fp32 quotient. numpy's float32 division is the IEEE quotient.
The operands cover the shared-step range [2^-40, 2^40], its edges, values
outside it, zeros, subnormals, infinities and NaN, and quotients placed
within a few ulps of a rounding midpoint, where a wrong rounding shows."""
import ctypes as C

import numpy as np
import pytest

from compute_rasterizer_amd import capi

pytestmark = pytest.mark.gpu


def device_divide(ctx, nx, ny, den):
    nx, ny, den = (np.ascontiguousarray(a, np.float32) for a in (nx, ny, den))
    out = np.empty(2 * den.size, np.float32)
    capi.check(capi.load().pcr_selftest_divide(ctx.h, capi.ptr(nx), capi.ptr(ny), capi.ptr(den), den.size,
                                               capi.ptr(out)))
    return out[0::2], out[1::2]


def same_bits(a, b):
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float32)
    both_nan = np.isnan(a) & np.isnan(b)
    return (a.view(np.uint32) == b.view(np.uint32)) | both_nan


def pow2_floats(rng, n, e_lo, e_hi):
    m = rng.uniform(1.0, 2.0, n)
    return np.ldexp(m, rng.integers(e_lo, e_hi + 1, n)).astype(np.float32)


def check(ctx, nx, ny, den):
    qx, qy = device_divide(ctx, nx, ny, den)
    with np.errstate(all="ignore"):
        ex = np.float32(nx) / np.float32(den)
        ey = np.float32(ny) / np.float32(den)
    okx, oky = same_bits(qx, ex), same_bits(qy, ey)
    bad = np.flatnonzero(~(okx & oky))
    assert bad.size == 0, (f"{bad.size} mismatches, first: num=({nx[bad[0]]!r}, {ny[bad[0]]!r}) "
                           f"den={den[bad[0]]!r} got=({qx[bad[0]]!r}, {qy[bad[0]]!r}) "
                           f"want=({ex[bad[0]]!r}, {ey[bad[0]]!r})")


def test_random_operands_across_and_beyond_the_range(ctx):
    rng = np.random.default_rng(11)
    n = 1 << 21
    den = pow2_floats(rng, n, -48, 48)
    nx = pow2_floats(rng, n, -48, 48) * rng.choice(np.float32([-1, 1]), n)
    ny = pow2_floats(rng, n, -48, 48) * rng.choice(np.float32([-1, 1]), n)
    check(ctx, nx, ny, den)


def test_clip_space_operands(ctx):
    # the shapes a frame produces: |x|, |y| up to a few times w, w in [0.01, 1e5]
    rng = np.random.default_rng(12)
    n = 1 << 21
    den = np.exp(rng.uniform(np.log(0.01), np.log(1e5), n)).astype(np.float32)
    nx = (den * rng.uniform(-3, 3, n)).astype(np.float32)
    ny = (den * rng.uniform(-3, 3, n)).astype(np.float32)
    check(ctx, nx, ny, den)


def test_near_rounding_midpoints(ctx):
    # numerators whose exact quotient lies within a few ulps of a midpoint
    # between two floats: num = RN(den * (q + ulp(q)/2)) nudged by -4..4 ulps
    rng = np.random.default_rng(13)
    n = 1 << 20
    den = pow2_floats(rng, n, -38, 38)
    q = pow2_floats(rng, n, -1, 0) * rng.choice(np.float32([-1, 1]), n)
    mid = q.astype(np.float64) + np.spacing(np.abs(q)).astype(np.float64) * 0.5 * np.sign(q)
    num = (den.astype(np.float64) * mid).astype(np.float32)
    bits = num.view(np.int32) + rng.integers(-4, 5, n).astype(np.int32)
    nx = bits.view(np.float32)
    ny = (den.astype(np.float64) * q.astype(np.float64)).astype(np.float32)
    check(ctx, nx, ny, den)


def test_range_edges_and_special_values(ctx):
    edge = np.float32([2.0 ** -40, 2.0 ** 40])
    near = np.concatenate([np.nextafter(edge, np.float32(0)), edge, np.nextafter(edge, np.float32(np.inf))])
    special = np.float32([0.0, -0.0, 1e-45, -1e-45, 1.1754942e-38, 3.4028235e38, -3.4028235e38,
                          np.inf, -np.inf, np.nan, 1.0, -1.0, 0.5, 3.0, 1e-12, 1e12])
    nums = np.concatenate([near, -near, special]).astype(np.float32)
    dens = np.concatenate([near, np.float32([1e-45, 1.1754942e-38, 1e-30, 1.0, 3.0, 1e30, 3.4028235e38, np.inf])])
    gx, gd = np.meshgrid(nums, dens, indexing="ij")
    gx = gx.ravel().astype(np.float32)
    gd = gd.ravel().astype(np.float32)
    gy = np.roll(gx, 7)
    check(ctx, gx, gy, gd)
