"""CPU: the oracle's LOD + hqs passes and the window-weighted resolve of
compute_loop_nodes_hqs (resolve.cs:92-158), pinned by closed forms and
invariants rather than by restating the formula."""
import numpy as np
import pytest

import compute_rasterizer_amd as pcr
from oracle import orc

NAN_DEPTH = 0xFFFFFFFF


def test_window0_equals_integer_division():
    # with count <= 1000, R/count is at least 1e-3 away from the next integer,
    # far beyond float rounding, so the float average truncates like integer division
    rng = np.random.default_rng(1)
    w, h = 64, 48
    cnt = rng.integers(0, 1000, w * h).astype(np.uint32)
    acc = np.zeros(4 * w * h, np.uint32)
    sums = []
    for k in range(3):
        # a multiple of the count plus a remainder below it: averages that are not integers
        s = rng.integers(0, 255, w * h) * cnt + rng.integers(0, np.maximum(cnt, 1))
        acc[k::4] = s.astype(np.uint32)
        sums.append(s)
    acc[3::4] = cnt
    depth = rng.integers(0x3F800000, 0x44000000, w * h).astype(np.uint32)
    out = orc.hqs_resolve_window(depth, acc, w, h, 0)
    c = np.maximum(cnt, 1).astype(np.int64)
    ref = np.where(cnt > 0, sums[0] // c | (sums[1] // c) << 8 | (sums[2] // c) << 16, 0)
    assert ((sums[0] % c) != 0).mean() > 0.9
    assert np.array_equal(out, ref)


@pytest.mark.parametrize("window", [1, 2, 3])
def test_uniform_region_keeps_its_colour(window):
    w, h = 20, 16
    acc = np.zeros(4 * w * h, np.uint32)
    acc[0::4], acc[1::4], acc[2::4], acc[3::4] = 37 * 3, 200 * 3, 5 * 3, 3
    depth = np.full(w * h, np.float32(12.5).view(np.uint32), np.uint32)
    out = orc.hqs_resolve_window(depth, acc, w, h, window)
    assert (out == (37 | 200 << 8 | 5 << 16)).all()


def test_deeper_neighbours_are_dropped():
    w, h = 9, 9
    acc = np.zeros(4 * w * h, np.uint32)
    acc[0::4], acc[3::4] = 250, 1                       # far pixels: red 250
    depth = np.full(w * h, np.float32(100.0).view(np.uint32), np.uint32)
    c = 4 * w + 4
    acc[4 * c:4 * c + 4] = [10, 20, 30, 1]              # the centre pixel is near
    depth[c] = np.float32(50.0).view(np.uint32)
    out = orc.hqs_resolve_window(depth, acc, w, h, 1)
    assert out[c] == (10 | 20 << 8 | 30 << 16)          # every neighbour is > 1.01 x 50 deep
    assert out[c + 2] == 250                            # away from the centre only far pixels remain


@pytest.mark.parametrize("window", [1, 2])
def test_edges_equal_a_padded_image(window):
    # neighbours outside the image contribute nothing: the same as empty
    # pixels (count 0, cleared depth) around it
    rng = np.random.default_rng(3)
    w, h, pad = 30, 20, 3
    cnt = rng.integers(0, 5, w * h).astype(np.uint32)
    acc = np.zeros(4 * w * h, np.uint32)
    for k in range(3):
        acc[k::4] = rng.integers(0, 256, w * h).astype(np.uint32) * cnt
    acc[3::4] = cnt
    depth = np.where(cnt > 0, rng.integers(0x41000000, 0x41200000, w * h), NAN_DEPTH).astype(np.uint32)
    W, H = w + 2 * pad, h + 2 * pad
    accp = np.zeros((H, W, 4), np.uint32)
    accp[pad:pad + h, pad:pad + w] = acc.reshape(h, w, 4)
    depp = np.full((H, W), NAN_DEPTH, np.uint32)
    depp[pad:pad + h, pad:pad + w] = depth.reshape(h, w)
    out = orc.hqs_resolve_window(depth, acc, w, h, window)
    outp = orc.hqs_resolve_window(depp.ravel(), accp.ravel(), W, H, window).reshape(H, W)
    assert np.array_equal(out.reshape(h, w), outp[pad:pad + h, pad:pad + w])


def test_lod_hqs_depth_is_the_lod_frame_depth():
    # pass 1 keeps the smallest depth per pixel over the same node cut and
    # decode as the LOD frame, whose keys carry that depth in their high bits
    xyz, rgba = orc.generate(pcr.gen_cfg(pcr.SCENE_TERRAIN, 60_000, seed=9))
    nodes, lo, med, hi, rgb = pcr.octree_layout_host(xyz, rgba, seed=2, max_node_points=3000)
    v = pcr.look_at((1000, -400, 600), (1000, 1000, 0), up=(0, 0, 1), near=1, far=5000, width=400, height=300)
    for rule in (pcr.LOD_SPACING, pcr.LOD_SIZE80):
        depth, acc = orc.hqs_nodes(nodes, lo, med, hi, rgb, v, 1, rule, 1.0)
        keys, flags = orc.render_nodes(nodes, lo, med, hi, v, 1, rule, 1.0)
        covered = keys != pcr.FB_CLEAR
        assert covered.sum() > 1000
        assert np.array_equal(depth[covered], (keys[covered] >> np.uint64(32)).astype(np.uint32))
        assert (depth[~covered] == NAN_DEPTH).all() and (acc.reshape(-1, 4)[~covered] == 0).all()
        assert (acc.reshape(-1, 4)[covered, 3] >= 1).all()
