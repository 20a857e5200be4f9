"""CPU: the oracle's eye-dome lighting resolve (compute_loop_las/resolve.cs:
148-191, which the reference carries switched off) and its exp.

The exp is pinned against the correctly rounded e^x; the shade against
closed forms on hand-built framebuffers (flat neighbourhoods keep their
colour, a pixel deeper than its neighbours darkens by exp(-(8d/9)*300*s),
background and out-of-image neighbours add nothing, alpha is dropped)."""
import numpy as np
import pytest

from oracle import orc

BG = 0x00443322
CLEAR = np.uint64(0xFFFFFFFFFFFFFFFF)


def key(depth, idx):
    return np.uint64((int(np.float32(depth).view(np.uint32)) << 32) | idx)


def ulps(a, b):
    return abs(int(np.float32(a).view(np.int32)) - int(np.float32(b).view(np.int32)))


def test_exp_is_within_one_ulp_of_correctly_rounded():
    xs = np.concatenate([np.linspace(-85.99, 0.0, 60001, dtype=np.float32),
                         -np.logspace(-9, 1.93, 3000).astype(np.float32)])
    worst = max(ulps(orc.edl_exp(x), np.float32(np.exp(np.float64(x)))) for x in xs)
    assert worst <= 1


def test_exp_special_values():
    assert orc.edl_exp(0.0) == 1.0 and orc.edl_exp(-0.0) == 1.0
    assert orc.edl_exp(-86.0) == 0.0 and orc.edl_exp(-1000.0) == 0.0
    # below the cut-off the true value times any 8-bit channel truncates to 0
    assert 255.0 * np.exp(-86.0) < 1.0
    assert orc.edl_exp(np.float32(np.log(0.5))) == pytest.approx(0.5, rel=1e-7)
    xs = np.linspace(-20.0, 0.0, 4001, dtype=np.float32)
    ys = np.array([orc.edl_exp(x) for x in xs])
    assert np.all(np.diff(ys) >= 0)   # monotone


def test_flat_neighbourhood_keeps_colour_without_alpha():
    w, h = 5, 4
    fb = np.array([key(12.5, i) for i in range(w * h)], np.uint64)
    rgba = (np.arange(w * h, dtype=np.uint32) * 0x010203 + 0xFF000000).astype(np.uint32)
    out = orc.resolve_edl(fb, w, h, rgba, BG, 0.0005)
    assert np.array_equal(out, rgba & 0x00FFFFFF)


@pytest.mark.parametrize("strength", [0.0005, 0.01])
def test_deeper_pixel_darkens_by_closed_form(strength):
    w, h, d, near = 3, 3, 40.0, 36.0
    fb = np.array([key(near, i) for i in range(9)], np.uint64)
    fb[4] = key(d, 4)                           # centre, 4 m behind its 8 neighbours
    rgba = np.full(9, 0x00C89664, np.uint32)    # R 100, G 150, B 200
    out = orc.resolve_edl(fb, w, h, rgba, BG, strength)
    shade = np.exp(-(8 * (d - near) / 9) * 300 * strength)
    for ch, c in ((0, 100), (8, 150), (16, 200)):
        got = (int(out[4]) >> ch) & 255
        assert abs(got - c * shade) <= 1.0   # truncation of c * shade
    assert int(out[4]) >> 24 == 0
    # the neighbours are nearer than the centre: nothing to subtract
    assert np.all(np.delete(out, 4) == 0x00C89664)


def test_background_and_image_edges_add_nothing():
    # a single point in an empty 4x3 image: its neighbours are background
    # (depth 1e8) or outside the image, so it keeps its colour
    w, h = 4, 3
    fb = np.full(w * h, CLEAR, np.uint64)
    fb[0] = key(5.0, 0)
    out = orc.resolve_edl(fb, w, h, np.array([0x11223344], np.uint32), BG, 0.01)
    assert out[0] == 0x00223344
    assert np.all(out[1:] == BG)
    # an index outside the cloud resolves to the background
    fb[1] = key(5.0, 7)
    assert orc.resolve_edl(fb, w, h, np.array([0x11223344], np.uint32), BG, 0.01)[1] == BG


def test_edge_pixel_sees_only_in_image_neighbours():
    # 1x3 column: the middle pixel is deeper than both in-image neighbours;
    # the 6 out-of-image neighbours count as background
    fb = np.array([key(10.0, 0), key(13.0, 1), key(11.0, 2)], np.uint64)
    out = orc.resolve_edl(fb, 1, 3, np.full(3, 0x000000FF, np.uint32), BG, 0.01)
    s = np.float32((13.0 - 10.0) + (13.0 - 11.0)) / np.float32(9.0)
    shade = orc.edl_exp(np.float32(np.float32(-s * np.float32(300.0)) * np.float32(0.01)))
    assert out[1] == int(np.float32(255.0) * np.float32(shade))
    assert out[0] == 0xFF and out[2] == 0xFF
