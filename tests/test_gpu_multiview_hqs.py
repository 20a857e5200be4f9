"""GPU: shared-decode multi-view hqs (pcr_multiview with PCR_MODE_HQS and
PCR_MODE_NODES_HQS), the N-view form of the two-eye VR methods
(compute_loop_las_hqs_vr/depth.cs:281-413, color.cs:294-513).

Each view's depth buffer, accumulators and image must equal the oracle's
single-view hqs of that view bit for bit, and the single-view GPU render of
the same view, over graph replays with eager work between them."""
import numpy as np
import pytest

import compute_rasterizer_amd as pcr
from oracle import orc

pytestmark = pytest.mark.gpu


def stereo_and_more(seed, eye_sep=3.0):
    """A stereo pair (two eyes eye_sep apart, same target offset) plus three
    views of other sizes."""
    rng = np.random.default_rng(seed)
    x, y = rng.uniform(600, 1400, 2)
    eye = np.array([x, y - 700.0, rng.uniform(250, 450)])
    tgt = np.array([x, y + 300.0, 40.0])
    right = np.array([eye_sep, 0.0, 0.0])
    views = [pcr.look_at(tuple(eye), tuple(tgt), up=(0, 0, 1), near=1.0, far=5000.0, width=1280, height=720),
             pcr.look_at(tuple(eye + right), tuple(tgt + right), up=(0, 0, 1), near=1.0, far=5000.0, width=1280,
                         height=720)]
    for w, h in [(640, 480), (256, 256), (1920, 1080)]:
        ang = rng.uniform(0, 2 * np.pi)
        e = (rng.uniform(300, 1700), rng.uniform(300, 1700), rng.uniform(150, 400))
        t = (e[0] + 300 * np.cos(ang), e[1] + 300 * np.sin(ang), 40.0)
        views.append(pcr.look_at(e, t, up=(0, 0, 1), near=1.0, far=5000.0, width=w, height=h))
    return views


@pytest.fixture(scope="module")
def city(ctx):
    cloud = ctx.generate(pcr.gen_cfg(pcr.SCENE_CITY, 1_500_000, seed=23, num_buildings=24))
    lay = cloud.download()
    yield cloud, lay
    cloud.close()


@pytest.fixture(scope="module")
def octree(ctx):
    xyz, rgba = orc.generate(pcr.gen_cfg(pcr.SCENE_TERRAIN, 300_000, seed=37))
    cloud = ctx.octree_host(xyz, rgba, seed=6, max_node_points=5000)
    lay = cloud.download()
    yield cloud, lay
    cloud.close()


def single_view(ctx, cloud, view, prm):
    fb = ctx.framebuffer(view.width, view.height, pcr.FB_HQS)
    ctx.render(cloud, view, prm, fb)
    ctx.resolve(cloud, prm, fb)
    d, a = fb.hqs()
    img = fb.rgba()
    fb.close()
    return d, a, img


@pytest.mark.parametrize("use_graph", [1, 0])
@pytest.mark.parametrize("tune", [(256, 40, 0, 0), (64, 160, 0, 0), (512, 80, 0, 0), (256, 40, 0, 2048),
                                  (128, 40, 0, 1024), (256, 40, 0, 4096)])
def test_multiview_hqs_equals_single_views(ctx, city, use_graph, tune):
    cloud, lay = city
    prm = pcr.params(pcr.MODE_HQS)
    views = stereo_and_more(1)
    fbs = [ctx.framebuffer(v.width, v.height, pcr.FB_HQS) for v in views]
    mv = pcr.MultiView(ctx, cloud, views, fbs, prm, tune=pcr.Tune(*tune), use_graph=bool(use_graph))
    for rnd in range(3):
        if rnd:
            views = stereo_and_more(1 + rnd)
            mv.set_views(views)
        mv.launch()
        ctx.synchronize()
        for k, (v, fb) in enumerate(zip(views, fbs)):
            d, a, img = single_view(ctx, cloud, v, prm)
            gd, ga = fb.hqs()
            assert np.array_equal(gd, d) and np.array_equal(ga, a), (rnd, k)
            assert np.array_equal(fb.rgba(), img), (rnd, k)
    # the oracle, for both eyes of the last round
    for k in (0, 1):
        rd, ra = orc.hqs_batched(lay["batches"], lay["lo"], lay["med"], lay["hi"], lay["rgba"], views[k])
        gd, ga = fbs[k].hqs()
        assert (ra[3::4] > 0).sum() > 1000
        assert np.array_equal(gd, rd) and np.array_equal(ga, ra)
        assert np.array_equal(fbs[k].rgba(), orc.hqs_resolve(ra, prm.background))
    assert not np.array_equal(fbs[0].rgba(), fbs[1].rgba())
    mv.close()
    for fb in fbs:
        fb.close()


def eyes(seed, n_pairs=1, ipd=0.065, sizes=((1280, 720), (1280, 720))):
    """n_pairs stereo pairs (eyes ipd apart) of the city from about 250 m, the
    two eyes at sizes[0] and sizes[1]."""
    rng = np.random.default_rng(seed)
    views = []
    for _ in range(n_pairs):
        x, y = rng.uniform(600, 1400, 2)
        eye = np.array([x, y - 700.0, rng.uniform(200, 300)])
        tgt = np.array([x, y + 300.0, 40.0])
        for off, (w, h) in zip((0.0, ipd), sizes):
            r = np.array([off, 0.0, 0.0])
            views.append(pcr.look_at(tuple(eye + r), tuple(tgt + r), up=(0, 0, 1), near=1.0, far=5000.0,
                                     width=w, height=h))
    return views


SAME, MIXED = ((1280, 720), (1280, 720)), ((1280, 720), (1920, 1080))


@pytest.mark.parametrize("n_pairs,tune,sizes,occl", [(1, (256, 40, 0, 2048), SAME, True),
                                                     (1, (128, 40, 0, 1024), SAME, True),
                                                     (2, (256, 40, 0, 4096), SAME, True),
                                                     (1, (256, 40, 0, 2048), MIXED, True),
                                                     (1, (256, 40, 0, 2048), SAME, False)])
def test_view_pairs_equal_oracle(ctx, city, n_pairs, tune, sizes, occl):
    """Two views of a batch at one level draw each hqs pass in one pass over
    the points (pass 1: two depth tiles; pass 2: the two halves of the pass-2
    tile); stereo pairs take those paths for most batches, also when the two
    views differ in size. Every view's depths, sums and image equal the
    oracle's, with and without the occlusion skip."""
    cloud, lay = city
    prm = pcr.params(pcr.MODE_HQS, occlusion_skip=occl)
    views = eyes(5 + n_pairs, n_pairs, sizes=sizes)
    fbs = [ctx.framebuffer(v.width, v.height, pcr.FB_HQS) for v in views]
    mv = pcr.MultiView(ctx, cloud, views, fbs, prm, tune=pcr.Tune(*tune))
    mv.launch()
    ctx.synchronize()
    for k, v in enumerate(views):
        rd, ra = orc.hqs_batched(lay["batches"], lay["lo"], lay["med"], lay["hi"], lay["rgba"], v)
        gd, ga = fbs[k].hqs()
        assert (ra[3::4] > 0).sum() > 1000
        assert np.array_equal(gd, rd), k
        assert np.array_equal(ga, ra), k
        assert np.array_equal(fbs[k].rgba(), orc.hqs_resolve(ra, prm.background)), k
    mv.close()
    for fb in fbs:
        fb.close()


@pytest.mark.parametrize("rule,window,tune", [(pcr.LOD_SPACING, 1, (128, 40, 0, 0)), (pcr.LOD_SIZE80, 2, (256, 40, 0, 2048)),
                                              (pcr.LOD_SPACING, 0, (128, 40, 0, 1024))])
def test_multiview_nodes_hqs_equals_oracle(ctx, octree, rule, window, tune):
    cloud, lay = octree
    prm = pcr.params(pcr.MODE_NODES_HQS, lod_rule=rule, lod_px=1.0, resolve_window=window)
    views = stereo_and_more(7)
    fbs = [ctx.framebuffer(v.width, v.height, pcr.FB_HQS) for v in views]
    mv = pcr.MultiView(ctx, cloud, views, fbs, prm, tune=pcr.Tune(*tune))
    for rnd in range(2):
        if rnd:
            views = stereo_and_more(8)
            mv.set_views(views)
        mv.launch()
        ctx.synchronize()
        for k, (v, fb) in enumerate(zip(views, fbs)):
            rd, ra = orc.hqs_nodes(lay["batches"], lay["lo"], lay["med"], lay["hi"], lay["rgba"], v, 1, rule, 1.0)
            gd, ga = fb.hqs()
            assert np.array_equal(gd, rd) and np.array_equal(ga, ra), (rnd, k)
            assert np.array_equal(fb.rgba(), orc.hqs_resolve_window(rd, ra, v.width, v.height, window)), (rnd, k)
            if k < 2:
                assert (ra[3::4] > 0).sum() > 1000
    mv.close()
    for fb in fbs:
        fb.close()


def test_multiview_hqs_autotune_and_refusals(ctx, city):
    cloud, _ = city
    views = stereo_and_more(3)[:2]
    fbs = [ctx.framebuffer(v.width, v.height, pcr.FB_HQS) for v in views]
    prm = pcr.params(pcr.MODE_HQS)
    best, ms = pcr.autotune(ctx, cloud, views, fbs, prm, iters=2)
    assert ms > 0
    with pytest.raises(pcr.PcrError, match="tile_px"):
        pcr.MultiView(ctx, cloud, views, fbs, prm, tune=pcr.Tune(512, 40, 0, 2048))
    with pytest.raises(pcr.PcrError, match="batch range"):
        pcr.MultiView(ctx, cloud, views, fbs, pcr.params(pcr.MODE_HQS, batch_begin=0, batch_end=1))
    keys = [ctx.framebuffer(v.width, v.height) for v in views]
    with pytest.raises(pcr.PcrError, match="HQS framebuffer"):
        pcr.MultiView(ctx, cloud, views, keys, prm)
    with pytest.raises(pcr.PcrError, match="KEY64 framebuffer"):
        pcr.MultiView(ctx, cloud, views, fbs, pcr.params(pcr.MODE_LAS2))
    for fb in fbs + keys:
        fb.close()
