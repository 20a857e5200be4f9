"""GPU: hqs accumulators wrap mod 2^32 channel by channel.

16.9M white points on one pixel push the R, G and B sums past 2^32
(255 x 16,900,000 = 4,309,500,000). The device adds each pixel's (R, G) and
(B, count) pairs with one 64-bit atomic and takes back the carry when the
low sum wraps (add_pairs); the result must be the four independent uint32
sums of hqs/color.cs's per-channel accumulation as the oracle computes it,
on the direct path, the tiled flush, the multi-view kernels and the
16 B/point path. The expected sums are also checked in closed form."""
import numpy as np
import pytest

import compute_rasterizer_amd as pcr
from oracle import orc

pytestmark = pytest.mark.gpu

N = 16_900_000
P = (10.0, 20.0, 5.0)


def view():
    return pcr.look_at((P[0], P[1] - 30.0, P[2] + 20.0), P, up=(0, 0, 1), near=1.0, far=1000.0, width=64, height=64)


@pytest.fixture(scope="module")
def pile(ctx):
    nb = (N + 10239) // 10240
    b = np.zeros(nb, pcr.BATCH_DTYPE)
    b["min"] = P   # zero-extent boxes: every point decodes to the box min
    b["max"] = P
    b["first_point"] = np.arange(nb, dtype=np.uint32) * 10240
    b["num_points"] = np.minimum(10240, N - np.arange(nb) * 10240)
    b["parent"] = -1
    z = np.zeros(N, np.uint32)
    rgba = np.full(N, 0x00FFFFFF, np.uint32)
    cloud = ctx.upload_batched(b, z, z, z, rgba)
    rd, ra = orc.hqs_batched(b, z, z, z, rgba, view())
    yield cloud, rd, ra
    cloud.close()


def check_wrapped(depth, acc, rd, ra):
    pix = int(np.flatnonzero(ra[3::4])[0])
    assert (ra[3::4] > 0).sum() == 1
    want = (255 * N) % 2**32
    assert want < 255 * N   # the channel sums did wrap
    assert list(ra[4 * pix:4 * pix + 4]) == [want, want, want, N]
    assert np.array_equal(depth, rd) and np.array_equal(acc, ra)


@pytest.mark.parametrize("tile", [0, 2048])
def test_single_view_wraps_like_uint32(ctx, pile, tile):
    cloud, rd, ra = pile
    v = view()
    fb = ctx.framebuffer(v.width, v.height, pcr.FB_HQS)
    ctx.render(cloud, v, pcr.params(pcr.MODE_HQS, tile_px=tile), fb)
    check_wrapped(*fb.hqs(), rd, ra)
    fb.close()


@pytest.mark.parametrize("tune", [(256, 40, 0, 0), (256, 40, 0, 2048)])
def test_multiview_wraps_like_uint32(ctx, pile, tune):
    cloud, rd, ra = pile
    views = [view(), view()]
    fbs = [ctx.framebuffer(64, 64, pcr.FB_HQS) for _ in views]
    mv = pcr.MultiView(ctx, cloud, views, fbs, pcr.params(pcr.MODE_HQS), tune=pcr.Tune(*tune))
    mv.launch()
    ctx.synchronize()
    for fb in fbs:
        check_wrapped(*fb.hqs(), rd, ra)
        fb.close()
    mv.close()


def test_points16_wraps_like_uint32(ctx, pile):
    _, rd, ra = pile
    xyz = np.tile(np.asarray(P, np.float32), (N, 1))
    rgba = np.full(N, 0x00FFFFFF, np.uint32)
    cloud = ctx.upload_points16(xyz, rgba)
    v = view()
    fb = ctx.framebuffer(v.width, v.height, pcr.FB_HQS)
    ctx.render(cloud, v, pcr.params(pcr.MODE_HQS16), fb)
    d, a = fb.hqs()
    rd16, ra16 = orc.hqs_points16(xyz, rgba, v)
    assert np.array_equal(ra16, ra)
    check_wrapped(d, a, rd16, ra16)
    fb.close()
    cloud.close()
