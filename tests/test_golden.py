"""Golden fixture tests (tests/golden/golden_small.npz, written by make_golden.py).

CPU: the oracle reproduces every recorded framebuffer and image bit for bit.
GPU: the device path reproduces them through the C ABI.
"""
import os

import numpy as np
import pytest

import compute_rasterizer_amd as pcr
from compute_rasterizer_amd.capi import BATCH_DTYPE, View
from oracle import orc

HERE = os.path.dirname(os.path.abspath(__file__))
G = np.load(os.path.join(HERE, "golden", "golden_small.npz"))
NVIEWS = 3


def view(k):
    return View.from_buffer_copy(G[f"view{k}"].tobytes())


def batches():
    return G["batches"].view(BATCH_DTYPE)


@pytest.mark.parametrize("k", range(NVIEWS))
def test_oracle_reproduces_golden(k):
    v = view(k)
    xyz, rgba = G["xyz"], G["rgba"]
    fb = orc.render_points16(xyz, rgba, v)
    assert np.array_equal(fb, G[f"plain_keys{k}"])
    assert np.array_equal(orc.resolve_color(fb, 0xFF000000), G[f"plain_rgba{k}"])
    fbb, levels = orc.render_batched(batches(), G["lo"], G["med"], G["hi"], v, 1)
    assert np.array_equal(fbb, G[f"batched_keys{k}"]) and np.array_equal(levels, G[f"batched_levels{k}"])
    assert np.array_equal(orc.resolve_index(fbb, rgba, 0x00443322), G[f"batched_rgba{k}"])
    d, a = orc.hqs_points16(xyz, rgba, v)
    assert np.array_equal(d, G[f"hqs_depth{k}"]) and np.array_equal(a, G[f"hqs_acc{k}"])
    assert np.array_equal(orc.hqs_resolve(a, 0x00443322), G[f"hqs_rgba{k}"])


def test_golden_inputs_regenerate():
    """The fixture's cloud and layout are what the seeded generator and encoder produce."""
    xyz, rgba = orc.generate(pcr.gen_cfg(pcr.SCENE_CUBE, 4096, seed=3))
    assert np.array_equal(xyz, G["xyz"]) and np.array_equal(rgba, G["rgba"])
    b, lo, med, hi = pcr.encode_batches_host(xyz, 1024)
    assert np.array_equal(b.view(np.uint8), G["batches"])
    assert np.array_equal(lo, G["lo"]) and np.array_equal(med, G["med"]) and np.array_equal(hi, G["hi"])


@pytest.mark.gpu
@pytest.mark.parametrize("k", range(NVIEWS))
def test_device_reproduces_golden(ctx, k):
    v = view(k)
    pts = ctx.upload_points16(G["xyz"], G["rgba"])
    fb = ctx.framebuffer(v.width, v.height)
    prm = pcr.params(pcr.MODE_PLAIN16)
    ctx.render(pts, v, prm, fb)
    ctx.resolve(pts, prm, fb)
    assert np.array_equal(fb.keys(), G[f"plain_keys{k}"])
    assert np.array_equal(fb.rgba(), G[f"plain_rgba{k}"])
    cloud = ctx.upload_batched(batches(), G["lo"], G["med"], G["hi"], G["rgba"])
    for mode, tile in ((pcr.MODE_LAS, 0), (pcr.MODE_LAS2, 0), (pcr.MODE_LAS2, 1024)):
        prm = pcr.params(mode, tile_px=tile)
        ctx.render(cloud, v, prm, fb)
        ctx.resolve(cloud, prm, fb)
        assert np.array_equal(fb.keys(), G[f"batched_keys{k}"])
        assert np.array_equal(fb.rgba(), G[f"batched_rgba{k}"])
    hfb = ctx.framebuffer(v.width, v.height, pcr.FB_HQS)
    prm = pcr.params(pcr.MODE_HQS16)
    ctx.render(pts, v, prm, hfb)
    ctx.resolve(pts, prm, hfb)
    d, a = hfb.hqs()
    assert np.array_equal(d, G[f"hqs_depth{k}"]) and np.array_equal(a, G[f"hqs_acc{k}"])
    assert np.array_equal(hfb.rgba(), G[f"hqs_rgba{k}"])
    for x in (hfb, fb, cloud, pts):
        x.close()
