"""Known-answer and closed-form pins of the oracle functions whose earlier
checks left room for plausible mistakes. None of these reads the golden
snapshot (which the oracle itself wrote); each case is chosen so that a moved
threshold, a rounding in place of a truncation, an extra low bit or a dropped
background rule fails it. CPU only.

  precision thresholds   compute_loop_las/render.cs:199-231, nodes/render.cs:180-212, :267-275
  decode levels 0 and 1  compute_loop_las2/render.cs:340-366, :402-424
  index resolve          compute_loop_las/resolve.cs:111-146 (colour gathered by point index)
  hqs resolve            compute_loop_las_hqs/resolve.cs:111-122 (uint division, alpha 0)
  LOD spacing cut        a node is refined when its projected spacing exceeds the LOD threshold
  LAS colour             modules/compute/LasLoaderSparse.cpp:184-198 (v < 256 ? v : v / 256)
  cube colours           the plain 16 B/pt workload: RGBA = uint8(255 x), uint8(255 y), uint8(255 z), 255
"""
import math
import struct

import numpy as np
import pytest

from compute_rasterizer_amd.capi import BATCH_DTYPE, FB_CLEAR, GenCfg, LOD_SIZE80, LOD_SPACING, SCENE_CUBE
from compute_rasterizer_amd.datasets import make_las
from oracle import orc

EPS = 0.01   # relative distance from a threshold: far above the float error of the projected size (~1e-6)


def _straight_view(w=1000, h=1000):
    return orc.look_at((0, 0, 0), (0, 0, -1), (0, 1, 0), fovy=90.0, near=1e-4, far=1e7, width=w, height=h)


def _cube_at(d, diag):
    """A cube of diagonal `diag` centred straight ahead at distance d."""
    hs = diag / math.sqrt(3.0) / 2.0
    return np.float32([-hs, -hs, -d - hs]), np.float32([hs, hs, -d + hs])


def _box_for_px(v, px, diag=1.0):
    """A box whose projected diagonal (render.cs:199-215) is `px` pixels:
    px = W/2 * P0 * diag / d for a box centred on the optical axis."""
    d = v.width * float(np.float32(v.proj[0])) * diag / (2.0 * px)
    return _cube_at(d, diag)


# ------------------------------------------------------- precision select (a4)
@pytest.mark.parametrize("first,threshold,below,above", [
    (100.0, 100.0, 4, 3), (100.0, 200.0, 3, 2), (100.0, 500.0, 2, 1), (100.0, 10000.0, 1, 0),
    (80.0, 80.0, 4, 3), (80.0, 100.0, 3, 3), (80.0, 200.0, 3, 2), (80.0, 500.0, 2, 1), (80.0, 10000.0, 1, 0),
])
def test_precision_level_on_both_sides_of_each_threshold(first, threshold, below, above):
    v = _straight_view()
    for px, want in ((threshold * (1 - EPS), below), (threshold * (1 + EPS), above)):
        mn, mx = _box_for_px(v, px)
        diag = float(np.float32(np.linalg.norm(mx.astype(np.float64) - mn)))
        got_px = orc.projected_px(mn, mx, diag, v.view_world, v.proj, v.width, v.height)
        assert abs(got_px - px) / px < 1e-4, (px, got_px)
        assert orc.precision_level(mn, mx, v, first) == want, (first, px)


@pytest.mark.parametrize("first,threshold,at", [
    (100.0, 100.0, 3), (100.0, 200.0, 2), (100.0, 500.0, 1), (100.0, 10000.0, 0), (80.0, 80.0, 3)])
def test_precision_level_exactly_at_threshold(first, threshold, at):
    """A flat box of x-extent L centred 5 units ahead projects its diagonal
    to W/2 * (L/5 + 1) - W/2 pixels; for these L every step is exact in
    float32 (P0 = 1, sqrt(dx^2) = |dx|), so the size equals the threshold and
    the strict `px < threshold` puts the box in the finer level."""
    v = _straight_view()
    assert float(np.float32(v.proj[0])) == 1.0
    L = np.float32(threshold / 100.0)
    mn, mx = np.float32([-L / 2, 0, -5]), np.float32([L / 2, 0, -5])
    assert orc.projected_px(mn, mx, L, v.view_world, v.proj, v.width, v.height) == threshold
    assert orc.precision_level(mn, mx, v, first) == at


def test_batched_render_levels_at_thresholds():
    """compute_loop_las uses the 100/200/500/10000 px thresholds for its
    batches (render.cs:219-229): the level the batched renderer reports on
    each side of each threshold."""
    v = _straight_view()
    cases = [(100.0, 4, 3), (200.0, 3, 2), (500.0, 2, 1), (10000.0, 1, 0)]
    boxes, want = [], []
    for t, below, above in cases:
        for px, lv in ((t * (1 - EPS), below), (t * (1 + EPS), above)):
            boxes.append(_box_for_px(v, px))
            want.append(lv)
    b = np.zeros(len(boxes), BATCH_DTYPE)
    for k, (mn, mx) in enumerate(boxes):
        b[k]["min"], b[k]["max"], b[k]["num_points"], b[k]["first_point"], b[k]["parent"] = mn, mx, 1, k, -1
    planes = np.zeros(len(boxes), np.uint32)
    _, levels = orc.render_batched(b, planes, planes, planes, v, 1)
    assert list(levels) == want


def test_size80_node_rule_on_both_sides():
    """The nodes method draws a node iff its projected diagonal is >= 80 px
    (nodes/render.cs:271-275: level >= 4 -> skip)."""
    v = _straight_view()
    for px, want in ((80.0 * (1 - EPS), 0), (80.0 * (1 + EPS), 1)):
        mn, mx = _box_for_px(v, px)
        n = np.zeros(1, BATCH_DTYPE)
        n["min"], n["max"], n["num_points"], n["parent"], n["spacing"] = mn, mx, 1, -1, 1.0
        assert orc.lod_select(n, v, 1, LOD_SIZE80, 1.0)[0] == want, px


# ------------------------------------------------------- LOD spacing cut (a12)
def _chain(d, spacing, depth=3):
    """root -> child -> grandchild, each the first octant of its parent, the
    root a cube of side 2 centred at distance d on the optical axis."""
    nodes = np.zeros(depth, BATCH_DTYPE)
    mn, mx = np.float32([-1, -1, -d - 1]), np.float32([1, 1, -d + 1])
    for k in range(depth):
        nodes[k]["min"], nodes[k]["max"] = mn, mx
        nodes[k]["num_points"], nodes[k]["parent"], nodes[k]["level"] = 1, k - 1, k
        nodes[k]["spacing"] = spacing / 2 ** k
        mx = np.float32((mn.astype(np.float64) + mx) / 2)
    return nodes


@pytest.mark.parametrize("lod_px", [1.0, 2.5, 10.0])
def test_lod_spacing_cut_closed_form(lod_px):
    """A child is drawn iff its parent's projected spacing W/2 * P0 * s / d
    exceeds lod_px (every ancestor must refine)."""
    v = _straight_view()
    P0 = float(np.float32(v.proj[0]))
    s = 0.01
    for ratio, want_child in ((1 - EPS, 0), (1 + EPS, 1)):
        d = v.width * P0 * s / (2.0 * lod_px * ratio)          # root's projected spacing = ratio * lod_px
        nodes = _chain(d, s)
        flags = orc.lod_select(nodes, v, 1, LOD_SPACING, lod_px)
        assert flags[0] == 1 and flags[1] == want_child, (lod_px, ratio, flags)
        # the child's own spacing projects to about ratio/2 * lod_px (half the
        # spacing, a little nearer), so the grandchild is never drawn
        assert flags[2] == 0


def test_lod_cut_needs_every_ancestor():
    v = _straight_view()
    P0 = float(np.float32(v.proj[0]))
    s = 0.01
    d = v.width * P0 * s / (2.0 * 4.0)                         # root spacing 4 px, child's about 2 px
    nodes = _chain(d, s)
    assert list(orc.lod_select(nodes, v, 1, LOD_SPACING, 1.0)) == [1, 1, 1]
    assert list(orc.lod_select(nodes, v, 1, LOD_SPACING, 3.0)) == [1, 1, 0]
    nodes[0]["spacing"] = s / 100                               # the root stops refining: nothing below it
    assert list(orc.lod_select(nodes, v, 1, LOD_SPACING, 1.0)) == [1, 0, 0]


# ------------------------------------------------------- decode, levels 0 and 1 (a5)
def _round_to_f32_int(X):
    """float(X) for non-negative integers X < 2^31, as exact integers: round
    to 24 significant bits, ties to even."""
    X = np.asarray(X, np.int64)
    bitlen = np.frexp(X.astype(np.float64))[1].astype(np.int64)        # exact for X < 2^53
    shift = np.maximum(bitlen - 24, 0)
    q = X >> shift
    rem = X & ((np.int64(1) << shift) - 1)
    half = (np.int64(1) << shift) >> 1
    up = (rem > half) | ((rem == half) & (half > 0) & ((q & 1) == 1))
    return (q + up) << shift


def _exact_decode(X, mn, mx):
    """fma(float(X), (mx - mn) * 2^-30, mn) rounded once to float32: the
    product of two 24-bit significands and the sum are exact in the 64-bit
    significand of long double for these boxes (asserted), so the final cast
    is the only rounding."""
    assert np.finfo(np.longdouble).nmant >= 63
    size = np.float32(np.float32(mx) - np.float32(mn))
    scale = np.float32(size * np.float32(2.0 ** -30))
    assert scale == np.float32(float(size) * 2.0 ** -30)                # a power-of-two scaling is exact
    F = _round_to_f32_int(X)
    # exactness: F < 2^31 and scale have 24-bit significands, so the product's
    # bits lie in [2^(Es-54), 2^(Es+1)) and mn's in [2^(Em-24), 2^Em), where
    # size < 2^Es and |mn| < 2^Em; the sum must fit 64 significant bits
    if mn != 0 and size != 0:
        Es, Em = math.frexp(float(size))[1], math.frexp(abs(float(mn)))[1]
        assert max(Es + 1, Em) + 1 - min(Es - 54, Em - 24) <= 64
    exact = F.astype(np.longdouble) * np.longdouble(scale) + np.longdouble(np.float32(mn))
    return exact.astype(np.float32)


BOXES = [([-3.25, 100.5, 0.0], [1.75, 612.5, 7.5]),
         ([1000.0, -50.0, 2.0], [1300.0, -30.0, 7.0]),
         ([-0.001, 0.25, -7777.0], [0.002, 0.251, -7000.0])]


def _perm(n, mult):
    return (np.arange(n, dtype=np.int64) * mult) % n                     # a bijection for odd mult


@pytest.mark.parametrize("box", range(len(BOXES)))
def test_decode_level1_exhaustive(box):
    """Level 1: X = lo << 20 | med << 10 for every (lo, med) pair on every axis;
    the fine plane is ignored (random bits there)."""
    mn, mx = (np.float32(v) for v in BOXES[box])
    n = 1 << 20
    pairs = [_perm(n, m) for m in (1, 7919, 104729)]                    # a different pair order per axis
    lo = np.zeros(n, np.uint32)
    med = np.zeros(n, np.uint32)
    for a, p in enumerate(pairs):
        lo |= ((p >> 10) << (10 * a)).astype(np.uint32)
        med |= ((p & 1023) << (10 * a)).astype(np.uint32)
    hi = np.random.default_rng(box).integers(0, 2 ** 30, n).astype(np.uint32)
    got = orc.decode_n(1, lo, med, hi, mn, mx)
    for a, p in enumerate(pairs):
        X = ((p >> 10) << 20) | ((p & 1023) << 10)
        want = _exact_decode(X, mn[a], mx[a])
        bad = np.flatnonzero(got[:, a] != want)
        assert bad.size == 0, (a, X[bad[:3]], got[bad[:3], a], want[bad[:3]])


@pytest.mark.parametrize("box", range(len(BOXES)))
def test_decode_level0_exhaustive_low_bits(box):
    """Level 0: X = lo << 20 | med << 10 | hi, every (med, hi) pair under
    coarse words on both sides of the 2^24 rounding boundary, on every axis."""
    mn, mx = (np.float32(v) for v in BOXES[box])
    n = 1 << 20
    for coarse in (0, 15, 16, 17, 1023):
        pairs = [_perm(n, m) for m in (1, 7919, 104729)]
        lo = np.zeros(n, np.uint32)
        med = np.zeros(n, np.uint32)
        hi = np.zeros(n, np.uint32)
        for a, p in enumerate(pairs):
            lo |= np.uint32(((coarse + 3 * a) % 1024) << (10 * a))
            med |= ((p >> 10) << (10 * a)).astype(np.uint32)
            hi |= ((p & 1023) << (10 * a)).astype(np.uint32)
        got = orc.decode_n(0, lo, med, hi, mn, mx)
        for a, p in enumerate(pairs):
            X = (np.int64((coarse + 3 * a) % 1024) << 20) | ((p >> 10) << 10) | (p & 1023)
            want = _exact_decode(X, mn[a], mx[a])
            bad = np.flatnonzero(got[:, a] != want)
            assert bad.size == 0, (coarse, a, X[bad[:3]], got[bad[:3], a], want[bad[:3]])


def test_decode_levels_2_to_4_ignore_the_fine_planes():
    mn, mx = np.float32([-3.25, 100.5, 7.0]), np.float32([1.75, 612.5, 7.5])
    rng = np.random.default_rng(4)
    lo = rng.integers(0, 2 ** 30, 4096).astype(np.uint32)
    med, hi = (rng.integers(0, 2 ** 30, 4096).astype(np.uint32) for _ in range(2))
    ref = orc.decode_n(2, lo, None, None, mn, mx)
    for level in (2, 3, 4):
        assert np.array_equal(orc.decode_n(level, lo, med, hi, mn, mx), ref)
    for a in range(3):
        X = ((lo >> (10 * a)) & 1023).astype(np.int64) << 20                 # the 10-bit value at 2^20 scale
        assert np.array_equal(ref[:, a], _exact_decode(X, mn[a], mx[a]))


def test_decode_known_answers():
    """Hand-computed values: box [0, 1024] per axis, so (mx - mn) * 2^-30 = 2^-20."""
    mn, mx = np.float32([0, 0, 0]), np.float32([1024, 1024, 1024])
    w = lambda x, y, z: x | y << 10 | z << 20                                # noqa: E731
    # level 1: X = 1 << 20 | 1 << 10 -> 1 + 2^-10; the fine word is ignored
    assert list(orc.decode(1, w(1, 0, 1023), w(1, 2, 0), w(1023, 1023, 1023), mn, mx)) == \
        [1 + 2 ** -10, 2 * 2 ** -10, 1023.0]
    # level 0: X = 1 << 20 | 0 | 512 -> 1 + 2^-11 (exact: 22 significant bits)
    assert list(orc.decode(0, w(1, 0, 0), 0, w(512, 3, 0), mn, mx)) == [1 + 2 ** -11, 3 * 2 ** -20, 0.0]
    # level 0 above 2^24: X = 1023 << 20 | 1023 << 10 | 1023 = 2^30 - 1 rounds to 2^30 -> 1024.0
    assert orc.decode(0, w(1023, 0, 0), w(1023, 0, 0), w(1023, 0, 0), mn, mx)[0] == 1024.0
    # X = 16 << 20 | 1 -> 2^24 + 1, a tie, rounds to even 2^24 -> 16.0
    assert orc.decode(0, w(16, 0, 0), 0, w(1, 0, 0), mn, mx)[0] == 16.0
    # X = 16 << 20 | 3 -> 2^24 + 3, a tie, rounds to even 2^24 + 4 -> 16 + 2^-18
    assert orc.decode(0, w(16, 0, 0), 0, w(3, 0, 0), mn, mx)[0] == 16.0 + 4 * 2 ** -20


# ------------------------------------------------------- index resolve (a7)
def test_index_resolve_closed_form():
    """out = rgba[key & 0xFFFFFFFF] with all 32 bits (alpha included) for every
    key other than the clear value; point index 0 is a point; no depth is
    capped; an index outside the cloud resolves to the background."""
    rng = np.random.default_rng(11)
    n, npix, bg = 5000, 40000, 0x00443322
    rgba = rng.integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32)
    rgba[0] = 0xA1B2C3D4
    depths = np.array([struct.unpack("<I", struct.pack("<f", f))[0]
                       for f in (0.5, 1.0, 999999.0, 1e6, 5e7, 3.0e38, -2.0, 0.0)] + [0x7F800000, 0xFFFFFFFF],
                      np.uint64)
    idx = rng.integers(0, n, npix).astype(np.uint64)
    idx[:64] = 0
    keys = (depths[rng.integers(0, depths.size, npix)] << np.uint64(32)) | idx
    keys[rng.random(npix) < 0.2] = FB_CLEAR
    keys[100:110] = (np.uint64(0x3F800000) << np.uint64(32)) | np.uint64(n)            # outside the cloud
    keys[110] = np.uint64(0xFFFFFFFF00000000)                                           # NaN depth, index 0
    keys[111] = np.uint64(0x00000000FFFFFFFF)                                           # depth 0, index 2^32-1
    out = orc.resolve_index(keys, rgba, bg)
    for i in range(npix):
        k = int(keys[i])
        want = bg if k == FB_CLEAR or (k & 0xFFFFFFFF) >= n else int(rgba[k & 0xFFFFFFFF])
        assert int(out[i]) == want, (i, hex(k))
    assert (out[idx == 0] == 0xA1B2C3D4).sum() >= 40 and out[110] == 0xA1B2C3D4


# ------------------------------------------------------- hqs resolve (a11)
def test_hqs_resolve_truncates_and_has_no_alpha():
    # hand-computed: 5/2, 7/2, 9/2 truncate to 2, 3, 4 (rounding would give 3, 4, 5 or 2, 4, 4)
    acc = np.array([5, 7, 9, 2, 254, 1, 0, 255, 255, 255, 255, 1, 0, 0, 0, 0, 1000, 999, 998, 4], np.uint32)
    out = orc.hqs_resolve(acc, 0x00443322)
    assert list(out) == [2 | 3 << 8 | 4 << 16, 0, 0xFFFFFF, 0x00443322, 250 | 249 << 8 | 249 << 16]


def test_hqs_resolve_integer_division_random():
    rng = np.random.default_rng(12)
    npix = 50000
    cnt = rng.integers(0, 5000, npix).astype(np.int64)
    cnt[:500] = 0
    sums = [rng.integers(0, 255 * cnt + 1) for _ in range(3)]               # not multiples of the count
    acc = np.zeros(4 * npix, np.uint32)
    for k in range(3):
        acc[k::4] = sums[k]
    acc[3::4] = cnt
    out = orc.hqs_resolve(acc, 0x00443322).astype(np.int64)
    c = np.maximum(cnt, 1)
    want = np.where(cnt > 0, sums[0] // c | (sums[1] // c) << 8 | (sums[2] // c) << 16, 0x00443322)
    assert np.array_equal(out, want)
    assert ((sums[0] % c) != 0).mean() > 0.9


# ------------------------------------------------------- LAS colour (f1)
LAS_COLOUR_KAT = [(0, 0), (1, 1), (128, 128), (254, 254), (255, 255), (256, 1), (257, 1), (300, 1), (511, 1),
                  (512, 2), (767, 2), (768, 3), (1000, 3), (65279, 254), (65280, 255), (65535, 255)]


def test_las_colour_16_to_8_bit_boundaries():
    n = len(LAS_COLOUR_KAT)
    data, *_ = make_las(n, point_format=2, seed=1)
    buf = bytearray(data)
    off, rl, rgb = struct.unpack_from("<I", buf, 96)[0], struct.unpack_from("<H", buf, 105)[0], 20
    vals = [v for v, _ in LAS_COLOUR_KAT]
    for i in range(n):
        struct.pack_into("<3H", buf, off + i * rl + rgb, vals[i], vals[n - 1 - i], vals[(i * 5) % n])
    info = orc.las_parse_header(bytes(buf))
    *_, rgba = orc.las_encode(bytes(buf), info)
    eight = [e for _, e in LAS_COLOUR_KAT]
    for i in range(n):
        assert int(rgba[i]) == eight[i] | eight[n - 1 - i] << 8 | eight[(i * 5) % n] << 16, i


# ------------------------------------------------------- cube workload colours (C1 input)
def test_cube_colours_are_uint8_of_255_xyz():
    xyz, rgba = orc.generate(GenCfg(SCENE_CUBE, 0, 5, 200000, 0, 0, 0, 0))
    assert xyz.dtype == np.float32 and (xyz >= 0).all() and (xyz < 1).all()
    ch = np.floor(np.float32(255.0) * xyz).astype(np.uint32)             # float32 product, truncated
    want = ch[:, 0] | ch[:, 1] << 8 | ch[:, 2] << 16 | np.uint32(0xFF000000)
    assert np.array_equal(rgba, want)
    # every 8-bit value occurs on every axis, 255 never (x < 1)
    for a in range(3):
        assert set(np.unique(ch[:, a])) == set(range(255))
