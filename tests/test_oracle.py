"""Pins the CPU oracle (oracle/pcr_oracle.c) against closed forms, published
vectors, brute force and invariants — each check is chosen so that a dropped
term, a swapped axis or a wrong rounding step fails it. CPU only.
"""
import math
import struct

import numpy as np
import pytest

from compute_rasterizer_amd.capi import BATCH_DTYPE, FB_CLEAR, GenCfg, View
from oracle import orc

M64 = (1 << 64) - 1
GOLDEN = 0x9E3779B97F4A7C15


def f32(x):
    return float(np.float32(x))


def fbits(x):
    return struct.unpack("<I", struct.pack("<f", x))[0]


# ----------------------------------------------------------------- PRNG
def test_splitmix64_published_vector():
    # Vigna's splitmix64.c, seed 1234567: the first five outputs of next()
    want = [6457827717110365317, 3203168211198807973, 9817491932198370423, 4593380528125082431,
            16408922859458223821]
    got = [orc.load().orc_mix64((1234567 + k * GOLDEN) & M64) for k in range(5)]
    assert got == want


def test_u01_closed_form():
    u = orc.load().orc_u01
    assert u(0) == 0.0
    assert u(1 << 40) == 2.0 ** -24
    assert u(M64) == 1.0 - 2.0 ** -24
    assert u(1 << 63) == 0.5


# ----------------------------------------------------------------- Morton
def _interleave_py(x, y, z):
    k = 0
    for b in range(21):
        k |= ((x >> b) & 1) << (3 * b + 2) | ((y >> b) & 1) << (3 * b + 1) | ((z >> b) & 1) << (3 * b)
    return k


def test_morton_bits():
    m = orc.load().orc_morton3
    assert m(1, 0, 0) == 4 and m(0, 1, 0) == 2 and m(0, 0, 1) == 1   # include/utils.h:26 convention
    assert m(2 ** 21 - 1, 2 ** 21 - 1, 2 ** 21 - 1) == 2 ** 63 - 1
    rng = np.random.default_rng(0)
    for x, y, z in rng.integers(0, 2 ** 21, (200, 3)):
        assert m(int(x), int(y), int(z)) == _interleave_py(int(x), int(y), int(z))


def test_morton_order_sorted_stable_permutation():
    rng = np.random.default_rng(1)
    xyz = rng.random((5000, 3), dtype=np.float32)
    xyz[100] = xyz[7]                 # a duplicate: stable order keeps 7 before 100
    perm, keys = orc.morton_order(xyz)
    assert np.array_equal(np.sort(perm), np.arange(5000))
    assert np.all(keys[1:] >= keys[:-1])
    assert list(perm).index(7) < list(perm).index(100)
    # independent key computation (python bit interleave, cube of the AABB)
    mn, mx = xyz.min(0), xyz.max(0)
    side = max(float(np.float32(mx[a] - mn[a])) for a in range(3))
    scale = 2097152.0 / side
    q = lambda v, m: min(int((float(v) - float(m)) * scale), 2097151)  # noqa: E731
    for i in (0, 17, 2500, 4999):
        j = perm[i]
        assert keys[i] == _interleave_py(q(xyz[j, 0], mn[0]), q(xyz[j, 1], mn[1]), q(xyz[j, 2], mn[2]))


# ----------------------------------------------------------------- encode / decode
def test_quantize30_closed_form():
    assert orc.quantize30(3.0, 3.0, 2.0) == 0
    assert orc.quantize30(5.0, 3.0, 2.0) == 2 ** 30 - 1          # clamp of 2^30
    assert orc.quantize30(4.0, 3.0, 2.0) == 2 ** 29
    assert orc.quantize30(7.0, 7.0, 0.0) == 0                    # zero extent
    for k in range(0, 1024, 37):
        assert orc.quantize30(10.0 + k / 1024.0, 10.0, 1.0) == k << 20


def test_encode_corners_and_header():
    xyz = np.array([[1, 2, 3], [5, 10, 4], [3, 6, 3.5], [1, 10, 3], [9, 9, 9]], np.float32)
    b, lo, med, hi = orc.encode_batches(xyz, 4)
    assert len(b) == 2
    assert b[0]["num_points"] == 4 and b[0]["first_point"] == 0
    assert b[1]["num_points"] == 1 and b[1]["first_point"] == 4
    assert list(b[0]["min"]) == [1, 2, 3] and list(b[0]["max"]) == [5, 10, 4]
    assert lo[0] == 0 and med[0] == 0 and hi[0] == 0                                    # batch min
    full = 1023 | 1023 << 10 | 1023 << 20
    assert lo[1] == full and med[1] == full and hi[1] == full                           # batch max (clamped)
    assert lo[2] == (512 | 512 << 10 | 512 << 20) and med[2] == 0 and hi[2] == 0       # exact midpoint
    assert lo[4] == 0 and med[4] == 0 and hi[4] == 0                                    # single point, zero extent
    assert b[1]["parent"] == -1


def test_decode_exhaustive_10bit():
    """Levels 2-4: x = fma(X, size/1024, min) for every 10-bit X, checked via an exact double sum."""
    mn, mx = np.float32([-3.25, 100.5, 7.0]), np.float32([1.75, 612.5, 7.5])
    for X in range(1024):
        w = X | X << 10 | X << 20
        got = orc.decode(3, w, 0, 0, mn, mx)
        for a in range(3):
            scale = np.float32(np.float32(mx[a] - mn[a]) * np.float32(2.0 ** -10))
            ref = np.float32(float(X) * float(scale) + float(mn[a]))   # exact in double -> one rounding
            assert got[a] == ref, (X, a)


@pytest.mark.parametrize("level,bits", [(0, 30), (1, 20), (2, 10)])
def test_encode_decode_roundtrip_error(level, bits):
    rng = np.random.default_rng(level)
    xyz = (rng.random((4096, 3)) * [300.0, 20.0, 5.0] + [1000.0, -50.0, 2.0]).astype(np.float32)
    b, lo, med, hi = orc.encode_batches(xyz, 4096)
    mn, mx = b[0]["min"], b[0]["max"]
    size = (mx - mn).astype(np.float64)
    for i in range(0, 4096, 7):
        p = orc.decode(level, int(lo[i]), int(med[i]), int(hi[i]), mn, mx)
        err = np.abs(p.astype(np.float64) - xyz[i])
        bound = size * 2.0 ** -bits + np.abs(xyz[i]) * 2.0 ** -22 + 1e-6
        assert np.all(err <= bound), (i, err, bound)


def test_decode_level0_uses_own_axis_bits():
    """Level 0 takes Y/Z low bits from their own fields (the las2 form, not X_12 for all axes)."""
    mn, mx = np.float32([0, 0, 0]), np.float32([1024, 1024, 1024])
    hi = 1 | (2 << 10) | (3 << 20)
    p = orc.decode(0, 0, 0, hi, mn, mx)
    assert list(p) == [f32(1 * 1024 * 2.0 ** -30), f32(2 * 1024 * 2.0 ** -30), f32(3 * 1024 * 2.0 ** -30)]


# ----------------------------------------------------------------- camera / projection
def test_look_at_closed_form():
    v = orc.look_at((0, 0, 5), (0, 0, 0), (0, 1, 0), fovy=90.0, near=1.0, far=101.0, width=200, height=100)
    t = np.array(v.transform).reshape(4, 4)      # column-major: t[col][row]
    assert t[0][0] == f32(0.5) and t[1][1] == f32(1.0)          # 1/(aspect*tan45), 1/tan45
    assert t[2][3] == -1.0 and t[3][3] == 5.0                   # w = 5 - z
    assert t[2][2] == f32(-102.0 / 100.0)
    vw = np.array(v.view_world).reshape(4, 4)
    assert vw[3][2] == -5.0 and vw[0][0] == 1.0


def test_project_known_pixels():
    W, H = 200, 100
    v = orc.look_at((0, 0, 5), (0, 0, 0), (0, 1, 0), fovy=90.0, near=1.0, far=101.0, width=W, height=H)
    t = np.array(v.transform, np.float32)
    assert orc.project(t, [0, 0, 0], W, H) == (100 + 50 * W, fbits(5.0))
    # ndc x = 0.5*x/(5-z): x = 9.9999 at z=0 -> ndc 0.99999 -> pixel 199
    assert orc.project(t, [9.9999, 0, 0], W, H)[0] % W == 199
    # ndc x exactly 1 -> truncated pixel == W -> rejected (edge decision)
    assert orc.project(t, [10.0, 0, 0], W, H) is None
    assert orc.project(t, [-10.0, 0, 0], W, H)[0] % W == 0
    assert orc.project(t, [0, 0, 6], W, H) is None               # behind the eye
    assert orc.project(t, [float("nan"), 0, 0], W, H) is None    # NaN is outside
    assert orc.project(t, [0, 5.0, 0], W, H) is None             # ndc y = 1 -> row H


def test_projected_px_and_levels_closed_form():
    """pixelSize = W*P0*r/(2d) for a box straight ahead; thresholds 100/200/500/10000."""
    W, H = 1000, 1000
    v = orc.look_at((0, 0, 0), (0, 0, -1), (0, 1, 0), fovy=90.0, near=0.1, far=1e6, width=W, height=H)
    P0 = float(np.float32(v.proj[0]))

    def box_at(d, r):  # a cube whose diagonal is r, centred at depth d
        h = r / math.sqrt(3) / 2
        return np.float32([-h, -h, -d - h]), np.float32([h, h, -d + h])

    for target, level in ((50, 4), (150, 3), (300, 2), (1000, 1), (20000, 0)):
        r = 10.0
        d = W * P0 * r / (2 * target)
        mn, mx = box_at(d, r)
        px = orc.projected_px(mn, mx, f32(np.linalg.norm(mx.astype(np.float64) - mn)), v.view_world, v.proj, W, H)
        assert abs(px - target) / target < 1e-3
        assert orc.precision_level(mn, mx, v) == level
    # nodes use 80 as the first threshold
    d = W * P0 * 10.0 / (2 * 90)
    mn, mx = box_at(d, 10.0)
    assert orc.precision_level(mn, mx, v, 100.0) == 4 and orc.precision_level(mn, mx, v, 80.0) == 3
    # monotone in distance
    lv = [orc.precision_level(*box_at(d, 10.0), v) for d in np.geomspace(0.3, 500, 40)]
    assert all(a <= b for a, b in zip(lv, lv[1:]))


# ----------------------------------------------------------------- frustum
def _planes64(m):
    m = np.asarray(m, np.float64)
    rows = []
    for s, ax in ((-1, 0), (1, 0), (1, 1), (-1, 1), (-1, 2), (1, 2)):
        p = np.array([m[4 * k + 3] + s * m[4 * k + ax] for k in range(4)])
        rows.append(p / np.linalg.norm(p[:3]))
    return rows


def test_frustum_cull_conservative_and_tight():
    v = orc.look_at((0, 0, 0), (0, 0, -1), (0, 1, 0), fovy=60.0, near=0.5, far=100.0, width=640, height=480)
    t = np.array(v.transform, np.float64).reshape(4, 4).T   # row-major
    planes = _planes64(v.transform)
    rng = np.random.default_rng(5)
    n_in = n_out = 0
    for _ in range(3000):
        c = rng.uniform([-80, -80, -150], [80, 80, 20])
        h = rng.uniform(0.1, 10, 3)
        mn, mx = (c - h).astype(np.float32), (c + h).astype(np.float32)
        vis = orc.frustum_visible(v.transform_frustum, mn, mx)
        # a sampled point of the box inside the clip volume forces "visible"
        pts = rng.uniform(mn, mx, (64, 3))
        clip = np.c_[pts, np.ones(64)] @ t.T
        inside = (clip[:, 3] > 0) & np.all(np.abs(clip[:, :3]) <= clip[:, 3:4] * (1 - 1e-4), axis=1)
        if inside.any():
            assert vis
            n_in += 1
        # the p-vertex of some plane clearly negative in double forces "culled"
        for p in planes:
            pv = np.where(p[:3] > 0, mx, mn).astype(np.float64)
            if p[:3] @ pv + p[3] < -1e-3:
                assert not vis
                n_out += 1
                break
    assert n_in > 100 and n_out > 100
    assert not orc.frustum_visible(v.transform_frustum, np.float32([-1, -1, 1]), np.float32([1, 1, 3]))  # behind
    assert orc.frustum_visible(v.transform_frustum, np.float32([-1, -1, -1]), np.float32([1, 1, 1]))     # contains eye


# ----------------------------------------------------------------- raster invariants
def _small_batched(n=3000, seed=3, bs=256):
    cfg = GenCfg(1, 0, seed, n, 200.0, 30.0, 0.5, 0.0)
    xyz, rgba = orc.generate(cfg)
    perm, _ = orc.morton_order(xyz)
    b, lo, med, hi = orc.encode_batches(xyz[perm], bs)
    return xyz[perm], rgba[perm], b, lo, med, hi


def _small_view(w=96, h=64):
    return orc.look_at((100, -60, 80), (100, 100, 10), (0, 0, 1), fovy=60, near=1, far=1000, width=w, height=h)


def test_batched_render_is_bruteforce_minimum():
    xyz, rgba, b, lo, med, hi = _small_batched()
    v = _small_view()
    fb, levels = orc.render_batched(b, lo, med, hi, v, 1)
    t = np.array(v.transform, np.float32)
    best = {}
    for k in range(len(b)):
        if levels[k] == 255:
            continue
        for i in range(b[k]["first_point"], b[k]["first_point"] + b[k]["num_points"]):
            p = orc.decode(int(levels[k]), int(lo[i]), int(med[i]), int(hi[i]), b[k]["min"], b[k]["max"])
            r = orc.project(t, p, v.width, v.height)
            if r:
                key = (r[1] << 32) | i
                best[r[0]] = min(best.get(r[0], FB_CLEAR), key)
    ref = np.full(v.width * v.height, FB_CLEAR, np.uint64)
    for pix, key in best.items():
        ref[pix] = key
    assert len(best) > 200
    assert np.array_equal(fb, ref)


def test_batched_render_order_invariant():
    """min is commutative: reversing the batch order gives the same bits."""
    xyz, rgba, b, lo, med, hi = _small_batched()
    v = _small_view()
    fb1, _ = orc.render_batched(b, lo, med, hi, v, 1)
    fb2, _ = orc.render_batched(np.ascontiguousarray(b[::-1]), lo, med, hi, v, 1)
    assert np.array_equal(fb1, fb2)
    fb3 = orc.render_batched_mt(b, lo, med, hi, v, 1, nthreads=4)
    assert np.array_equal(fb1, fb3)


def test_points16_merge_of_halves():
    cfg = GenCfg(0, 0, 9, 20000, 0, 0, 0, 0)
    xyz, rgba = orc.generate(cfg)
    v = orc.look_at((0.5, 0.5, 2.5), (0.5, 0.5, 0.5), width=128, height=72)
    full = orc.render_points16(xyz, rgba, v)
    a = orc.render_points16(xyz[:7000], rgba[:7000], v)
    bb = orc.render_points16(xyz[7000:], rgba[7000:], v)
    assert np.array_equal(full, np.minimum(a, bb))
    img = orc.resolve_color(full, 0xFF000000)
    drawn = full != FB_CLEAR
    assert np.all(img[~drawn] == 0xFF000000)
    assert np.all(img[drawn] == (full[drawn] & 0xFFFFFFFF).astype(np.uint32))
    assert np.all(img[drawn] >> 24 == 255)


def test_cull_changes_only_out_of_volume_points():
    """Culling removes whole batches; every pixel it changes had a point beyond near/far."""
    xyz, rgba, b, lo, med, hi = _small_batched(n=4000, seed=8)
    v = orc.look_at((100, -60, 80), (100, 100, 10), (0, 0, 1), fovy=60, near=1, far=120, width=96, height=64)
    on, lv = orc.render_batched(b, lo, med, hi, v, 1)
    off, _ = orc.render_batched(b, lo, med, hi, v, 0)
    assert (lv == 255).any()
    diff = on != off
    assert np.all(off[diff] <= on[diff])


# ----------------------------------------------------------------- hqs
def test_hqs_bruteforce_counts_and_sums():
    cfg = GenCfg(0, 0, 4, 6000, 0, 0, 0, 0)
    xyz, rgba = orc.generate(cfg)
    v = orc.look_at((0.5, 0.5, 2.5), (0.5, 0.5, 0.5), width=40, height=30)
    depth, acc = orc.hqs_points16(xyz, rgba, v)
    t = np.array(v.transform, np.float32)
    hits = {}
    for i in range(len(xyz)):
        r = orc.project(t, xyz[i], 40, 30)
        if r:
            hits.setdefault(r[0], []).append((r[1], rgba[i]))
    for pix, lst in hits.items():
        dmin = min(d for d, _ in lst)
        assert depth[pix] == dmin
        lim = np.float32(struct.unpack("<f", struct.pack("<I", dmin))[0]) * np.float32(1.01)
        sel = [c for d, c in lst if struct.unpack("<f", struct.pack("<I", d))[0] <= lim]
        assert acc[4 * pix + 3] == len(sel)
        assert acc[4 * pix + 0] == sum(int(c) & 255 for c in sel)
        assert acc[4 * pix + 2] == sum((int(c) >> 16) & 255 for c in sel)
    out = orc.hqs_resolve(acc, 0x00443322)
    assert np.all(out[acc[3::4] == 0] == 0x00443322)


def test_hqs_single_point_pixels_equal_plain_resolve():
    cfg = GenCfg(0, 0, 4, 50, 0, 0, 0, 0)
    xyz, rgba = orc.generate(cfg)
    v = orc.look_at((0.5, 0.5, 2.5), (0.5, 0.5, 0.5), width=640, height=480)
    depth, acc = orc.hqs_points16(xyz, rgba, v)
    plain = orc.resolve_color(orc.render_points16(xyz, rgba, v), 0x00443322)
    single = acc[3::4] == 1
    assert single.sum() >= 45
    # the hqs image carries r | g << 8 | b << 16 with alpha 0 (hqs/resolve.cs:119)
    assert np.array_equal(orc.hqs_resolve(acc, 0x00443322)[single], plain[single] & 0x00FFFFFF)


# ----------------------------------------------------------------- LOD
def _toy_tree():
    """root -> 8 children -> 64 grandchildren over [0,64]^3, spacing halves per level."""
    nodes = []

    def add(mn, size, level, parent):
        n = np.zeros(1, BATCH_DTYPE)[0]
        n["min"] = mn
        n["max"] = np.float32(mn) + size
        n["num_points"] = 4
        n["first_point"] = 4 * len(nodes)
        n["level"] = level
        n["parent"] = parent
        n["spacing"] = 64.0 / 128 / 2 ** level
        nodes.append(n)
        return len(nodes) - 1

    root = add([0, 0, 0], 64.0, 0, -1)
    kids = [add([x, y, z], 32.0, 1, root) for x in (0, 32) for y in (0, 32) for z in (0, 32)]
    for k in kids:
        base = nodes[k]["min"]
        for dx in (0, 16):
            for dy in (0, 16):
                for dz in (0, 16):
                    add(base + np.float32([dx, dy, dz]), 16.0, 2, k)
    return np.array(nodes, dtype=BATCH_DTYPE)


def test_lod_top_down_cut_and_extremes():
    nodes = _toy_tree()
    far = orc.look_at((32, 32, 5000), (32, 32, 0), (0, 1, 0), far=1e5, width=512, height=512)
    sel = orc.lod_select(nodes, far, 1, 0, 1.0)
    assert sel[0] == 1 and sel[1:].sum() == 0                  # far away: root only
    near = orc.look_at((32, 32, 70), (32, 32, 0), (0, 1, 0), far=1e5, width=2048, height=2048)
    sel = orc.lod_select(nodes, near, 1, 0, 1.0)
    assert sel.sum() > 9
    for i, n in enumerate(nodes):
        if sel[i] and n["parent"] >= 0:
            assert sel[n["parent"]]                            # top-down cut
    ref80 = orc.lod_select(nodes, near, 1, 1, 1.0)
    assert ref80.dtype == np.uint8
