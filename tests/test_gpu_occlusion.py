"""GPU: the occlusion skip of the tiled kernels and the frame counters.

A batch whose nearest possible depth lies behind every depth already held
over its screen rectangle is skipped (rect_open). The skip must never change
a bit: frames with buildings in front of other buildings and terrain equal
the oracle; a frame rendered a second time over its own result (no clear)
skips the batches hidden behind it and leaves the framebuffer unchanged; hqs
depth and accumulators equal the oracle's with the pass-2 skip at work.

The counters follow the reference's debug values (render.cs:73-81):
pixels_visible is the number of covered pixels (closed form from the keys),
and atomics_issued lies between that count and the points drawn.
"""
import numpy as np
import pytest

import compute_rasterizer_amd as pcr
from oracle import orc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def city(ctx):
    cfg = pcr.gen_cfg(pcr.SCENE_CITY, 3_000_000, seed=21, extent=400.0, amplitude=20.0, num_buildings=40,
                      terrain_frac=0.3)
    cloud = ctx.generate(cfg)
    yield cloud, cloud.download()
    cloud.close()


def _views(w=800, h=450):
    # low cameras looking across the blocks, so buildings hide buildings and terrain
    return [pcr.look_at((x, y, 35.0), (tx, ty, 10.0), up=(0, 0, 1), near=0.5, far=2000.0, width=w, height=h)
            for (x, y, tx, ty) in ((20.0, 30.0, 300.0, 350.0), (380.0, 200.0, 50.0, 180.0),
                                   (200.0, 10.0, 210.0, 390.0))]


def test_occluded_frames_equal_oracle(ctx, city):
    cloud, lay = city
    prm = pcr.params(pcr.MODE_LAS2, tile_px=2048, debug_counters=True)
    skipped = 0
    for k, v in enumerate(_views()):
        fb = ctx.framebuffer(v.width, v.height)
        ctx.render(cloud, v, prm, fb)
        ctx.resolve(cloud, prm, fb)
        st = ctx.stats()
        keys = fb.keys()
        ref, _ = orc.render_batched(lay["batches"], lay["lo"], lay["med"], lay["hi"], v, 1)
        print(f"view {k}: drawn {st.points_drawn} occluded {st.points_occluded} atomics {st.atomics_issued}",
              flush=True)
        assert (ref != pcr.FB_CLEAR).sum() > 10000
        assert np.array_equal(keys, ref), (k, int((keys != ref).sum()))
        assert np.array_equal(fb.rgba(), orc.resolve_index(ref, lay["rgba"], prm.background))
        covered = int((keys != pcr.FB_CLEAR).sum())
        assert st.pixels_visible == covered
        assert covered <= st.atomics_issued <= st.points_drawn
        skipped += st.points_occluded
        fb.close()
    assert skipped > 0


def test_second_render_over_its_own_frame_skips_and_changes_nothing(ctx, city):
    cloud, lay = city
    v = _views()[0]
    fb = ctx.framebuffer(v.width, v.height)
    ctx.render(cloud, v, pcr.params(pcr.MODE_LAS2, tile_px=2048), fb)
    first = fb.keys()
    st1 = ctx.stats()
    ctx.render(cloud, v, pcr.params(pcr.MODE_LAS2, tile_px=2048, clear_before=0), fb)
    st2 = ctx.stats()
    assert np.array_equal(fb.keys(), first)
    # over the finished frame every batch behind it is skipped, and nothing
    # needs an atomic
    assert st2.points_occluded > st1.points_occluded
    assert st2.points_occluded > 0
    assert st2.atomics_issued == 0
    assert st2.pixels_visible == 0          # counted only with debug_counters
    fb.close()


def test_skip_off_gives_the_same_frame(ctx, city):
    cloud, lay = city
    v = _views()[1]
    fb = ctx.framebuffer(v.width, v.height)
    ctx.render(cloud, v, pcr.params(pcr.MODE_LAS2, tile_px=2048), fb)
    on = fb.keys()
    ctx.render(cloud, v, pcr.params(pcr.MODE_LAS2, tile_px=2048, occlusion_skip=False), fb)
    assert ctx.stats().points_occluded == 0
    assert np.array_equal(fb.keys(), on)
    fb.close()


def test_hqs_with_skips_equals_oracle(ctx, city):
    cloud, lay = city
    prm = pcr.params(pcr.MODE_HQS, tile_px=2048)
    skipped = 0
    for k, v in enumerate(_views(480, 270)):
        fb = ctx.framebuffer(v.width, v.height, pcr.FB_HQS)
        ctx.render(cloud, v, prm, fb)
        ctx.resolve(cloud, prm, fb)
        st = ctx.stats()
        d, a = fb.hqs()
        rd, ra = orc.hqs_batched(lay["batches"], lay["lo"], lay["med"], lay["hi"], lay["rgba"], v, 1)
        assert np.array_equal(d, rd), (k, int((d != rd).sum()))
        assert np.array_equal(a, ra), (k, int((a != ra).sum()))
        assert np.array_equal(fb.rgba(), orc.hqs_resolve(ra, prm.background))
        skipped += st.points_occluded
        fb.close()
    assert skipped > 0


def test_multiview_with_skips_equals_single_views(ctx, city):
    cloud, lay = city
    views = _views(640, 360)
    prm = pcr.params(pcr.MODE_LAS2, tile_px=2048)
    fbs = [ctx.framebuffer(v.width, v.height) for v in views]
    mv = pcr.MultiView(ctx, cloud, views, fbs, prm, tune=pcr.Tune(256, 40, 0, 2048))
    mv.launch()
    ctx.synchronize()
    for k, v in enumerate(views):
        ref, _ = orc.render_batched(lay["batches"], lay["lo"], lay["med"], lay["hi"], v, 1)
        assert np.array_equal(fbs[k].keys(), ref), k
    mv.close()
    for fb in fbs:
        fb.close()
