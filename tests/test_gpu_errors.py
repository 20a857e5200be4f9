"""GPU: invalid inputs fail with a status code and a message (the reference
skipped silently), and repeated frames give identical bits (atomic min is
order independent, so any schedule must produce the same framebuffer)."""
import numpy as np
import pytest

import compute_rasterizer_amd as pcr
from compute_rasterizer_amd import capi
from oracle import orc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small(ctx):
    xyz, rgba = orc.generate(pcr.gen_cfg(pcr.SCENE_TERRAIN, 20_000, seed=8))
    b, lo, med, hi = pcr.encode_batches_host(xyz, 1024)
    return xyz, rgba, b, lo, med, hi


def expect_invalid(fn, text):
    with pytest.raises(pcr.PcrError) as e:
        fn()
    assert e.value.code == capi.PCR_E_INVALID
    assert text in str(e.value)


def test_upload_validation(ctx, small):
    xyz, rgba, b, lo, med, hi = small
    bad = b.copy()
    bad[-1]["num_points"] += 64
    expect_invalid(lambda: ctx.upload_batched(bad, lo, med, hi, rgba), "exceeds")
    bad = b.copy()
    bad[2]["parent"] = 5
    expect_invalid(lambda: ctx.upload_batched(bad, lo, med, hi, rgba), "parent")


def test_empty_and_oversized_inputs(ctx, small):
    """Clouds need points, framebuffers a size within PCR_MAX_DIM, and an
    empty batch range away from batch 0 draws nothing."""
    xyz, rgba, b, lo, med, hi = small
    z = np.zeros(0, np.uint32)
    expect_invalid(lambda: ctx.upload_batched(b[:0], z, z, z, z), "bad sizes")
    expect_invalid(lambda: ctx.upload_points16(np.zeros((0, 3), np.float32), z), "no points")
    from compute_rasterizer_amd.datasets import make_las
    expect_invalid(lambda: ctx.from_las([make_las(0, seed=2)[0]]), "0 points")
    for w, h in ((32769, 8), (8, 32769), (0, 8), (8, -1)):
        expect_invalid(lambda: ctx.framebuffer(w, h), "bad size")
    cloud = ctx.upload_batched(b, lo, med, hi, rgba)
    v = pcr.look_at((1000, -200, 400), (1000, 1000, 50), up=(0, 0, 1), near=1, far=5000, width=320, height=200)
    fb = ctx.framebuffer(320, 200)
    for mode in (pcr.MODE_LAS, pcr.MODE_LAS2):
        ctx.render(cloud, v, pcr.params(mode, batch_begin=3, batch_end=3), fb)
        assert (fb.keys() == pcr.FB_CLEAR).all()
    hfb = ctx.framebuffer(320, 200, pcr.FB_HQS)
    ctx.render(cloud, v, pcr.params(pcr.MODE_HQS, batch_begin=5, batch_end=5), hfb)
    d, a = hfb.hqs()
    assert (d == 0xFFFFFFFF).all() and (a == 0).all()
    fb.close()
    hfb.close()
    cloud.close()


def test_render_validation(ctx, small):
    xyz, rgba, b, lo, med, hi = small
    cloud = ctx.upload_batched(b, lo, med, hi, rgba)
    pts = ctx.upload_points16(xyz, rgba)
    v = pcr.look_at((1000, -200, 400), (1000, 1000, 50), up=(0, 0, 1), near=1, far=5000, width=320, height=200)
    fb = ctx.framebuffer(320, 200)
    hfb = ctx.framebuffer(320, 200, pcr.FB_HQS)
    small_fb = ctx.framebuffer(100, 100)
    expect_invalid(lambda: ctx.render(cloud, v, pcr.params(pcr.MODE_LAS2), small_fb), "differs from framebuffer")
    expect_invalid(lambda: ctx.render(cloud, v, pcr.params(pcr.MODE_HQS), fb), "HQS framebuffer")
    expect_invalid(lambda: ctx.render(cloud, v, pcr.params(pcr.MODE_LAS2), hfb), "KEY64 framebuffer")
    expect_invalid(lambda: ctx.render(pts, v, pcr.params(pcr.MODE_LAS2), fb), "no batched layout")
    expect_invalid(lambda: ctx.render(cloud, v, pcr.params(pcr.MODE_PLAIN16), fb), "16 B/pt")
    expect_invalid(lambda: ctx.render(cloud, v, pcr.params(pcr.MODE_LAS2, tile_px=3000), fb), "tile_px")
    expect_invalid(lambda: ctx.render(cloud, v, pcr.params(pcr.MODE_LAS2, block_size=96), fb), "block_size")
    expect_invalid(lambda: ctx.render(cloud, v, pcr.params(7), fb), "unknown mode")
    # after the errors the context still renders correctly
    prm = pcr.params(pcr.MODE_LAS2)
    ctx.render(cloud, v, prm, fb)
    ref, _ = orc.render_batched(b, lo, med, hi, v, 1)
    assert np.array_equal(fb.keys(), ref)


def test_multiview_validation(ctx, small):
    xyz, rgba, b, lo, med, hi = small
    cloud = ctx.upload_batched(b, lo, med, hi, rgba)
    v = pcr.look_at((1000, -200, 400), (1000, 1000, 50), up=(0, 0, 1), near=1, far=5000, width=64, height=64)
    fbs = [ctx.framebuffer(64, 64) for _ in range(65)]
    expect_invalid(lambda: pcr.MultiView(ctx, cloud, [v] * 65, fbs, pcr.params(pcr.MODE_LAS2)), "views")
    expect_invalid(lambda: pcr.MultiView(ctx, cloud, [v], fbs[:1], pcr.params(pcr.MODE_HQS16)), "mode")
    expect_invalid(lambda: pcr.MultiView(ctx, cloud, [v], fbs[:1], pcr.params(pcr.MODE_HQS)), "HQS framebuffer")
    expect_invalid(lambda: pcr.MultiView(ctx, cloud, [v], fbs[:1], pcr.params(pcr.MODE_LAS2),
                                         tune=pcr.Tune(512, 40, 0, 2048)), "tile_px")
    mv = pcr.MultiView(ctx, cloud, [v, v], fbs[:2], pcr.params(pcr.MODE_LAS2))
    v2 = pcr.look_at((0, 0, 10), (1, 1, 0), width=32, height=32)
    expect_invalid(lambda: mv.set_views([v2, v2]), "changed size")
    mv.close()


@pytest.mark.parametrize("tile", [0, 2048])
def test_repeated_frames_identical(ctx, tile):
    cloud = ctx.generate(pcr.gen_cfg(pcr.SCENE_CITY, 3_000_000, seed=31, num_buildings=32))
    v = pcr.look_at((1000, 200, 300), (1000, 1000, 20), up=(0, 0, 1), near=1, far=5000, width=1280, height=720)
    fb = ctx.framebuffer(1280, 720)
    prm = pcr.params(pcr.MODE_LAS2, tile_px=tile)
    first = None
    for _ in range(5):
        ctx.render(cloud, v, prm, fb)
        ctx.resolve(cloud, prm, fb)
        k = fb.keys()
        if first is None:
            first = k
        assert np.array_equal(k, first)
    assert (first != pcr.FB_CLEAR).sum() > 10_000
    cloud.close()


def test_frame_times_levels(ctx, small):
    """time_stages 1 records every stage; 2 only the raster, with the frame
    period taken from raster end to raster end; other values are refused."""
    xyz, rgba, b, lo, med, hi = small
    cloud = ctx.upload_batched(b, lo, med, hi, rgba)
    v = pcr.look_at((1000, -200, 400), (1000, 1000, 50), up=(0, 0, 1), near=1, far=5000, width=320, height=200)
    fb = ctx.framebuffer(320, 200)
    expect_invalid(lambda: ctx.render(cloud, v, pcr.params(pcr.MODE_LAS2, time_stages=3), fb), "time_stages")
    for level in (1, 2):
        prm = pcr.params(pcr.MODE_LAS2, time_stages=level)
        for _ in range(4):
            ctx.render(cloud, v, prm, fb)
            ctx.resolve(cloud, prm, fb)
        ft = ctx.frame_times(3)
        assert ft.shape == (3, 5)
        assert np.all(ft[:, 2] > 0)
        assert np.all(ft[:, 4] >= ft[:, 2])
        if level == 1:
            assert np.all(ft[:, [0, 1, 3]] > 0)
        else:
            assert np.all(ft[:, [0, 1, 3]] == -1)
    fb.close()
    cloud.close()
