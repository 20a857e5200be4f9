"""GPU: the eye-dome lighting resolve (params.edl_strength, the shading that
compute_loop_las/resolve.cs:148-191 carries switched off) equals the oracle's
orc_resolve_edl bit for bit, single-view and multi-view, and is refused for
modes without an index payload."""
import numpy as np
import pytest

import compute_rasterizer_amd as pcr
from oracle import orc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def city(ctx):
    cloud = ctx.generate(pcr.gen_cfg(pcr.SCENE_CITY, 1_000_000, seed=29, num_buildings=16))
    lay = cloud.download()
    yield cloud, lay
    cloud.close()


def views():
    return [pcr.look_at((900, 300, 260), (1000, 1000, 40), up=(0, 0, 1), near=1.0, far=5000.0, width=640, height=480),
            pcr.look_at((300, 900, 120), (1200, 1100, 60), up=(0, 0, 1), near=1.0, far=5000.0, width=331, height=197)]


@pytest.mark.parametrize("mode,strength", [(pcr.MODE_LAS2, 0.0005), (pcr.MODE_LAS2, 0.01), (pcr.MODE_LAS, 0.05)])
def test_edl_resolve_matches_oracle(ctx, city, mode, strength):
    cloud, lay = city
    for v in views():
        fb = ctx.framebuffer(v.width, v.height)
        prm = pcr.params(mode, edl_strength=strength)
        ctx.render(cloud, v, prm, fb)
        ctx.resolve(cloud, prm, fb)
        keys = fb.keys()
        ref = orc.resolve_edl(keys, v.width, v.height, lay["rgba"], prm.background, strength)
        img = fb.rgba()
        assert np.array_equal(img, ref)
        plain = orc.resolve_index(keys, lay["rgba"], prm.background)
        drawn = keys != pcr.FB_CLEAR
        assert drawn.sum() > 1000
        # shading only darkens, and it darkens somewhere at the depth edges
        assert np.all((img[drawn] & 255) <= (plain[drawn] & 255))
        assert np.any((img[drawn] & 0xFFFFFF) != (plain[drawn] & 0xFFFFFF))
        fb.close()


def test_edl_multiview_matches_oracle(ctx, city):
    cloud, lay = city
    vs = views()
    fbs = [ctx.framebuffer(v.width, v.height) for v in vs]
    prm = pcr.params(pcr.MODE_LAS2, edl_strength=0.01)
    mv = pcr.MultiView(ctx, cloud, vs, fbs, prm, tune=pcr.Tune(128, 40, 0, 1024))
    for _ in range(2):
        mv.launch()
        ctx.synchronize()
        for v, fb in zip(vs, fbs):
            ref = orc.resolve_edl(fb.keys(), v.width, v.height, lay["rgba"], prm.background, 0.01)
            assert np.array_equal(fb.rgba(), ref)
    mv.close()
    for fb in fbs:
        fb.close()


def test_edl_refused_without_index_payload(ctx, city):
    cloud, _ = city
    v = views()[0]
    fb = ctx.framebuffer(v.width, v.height, pcr.FB_HQS)
    with pytest.raises(pcr.PcrError, match="edl_strength"):
        ctx.resolve(cloud, pcr.params(pcr.MODE_HQS, edl_strength=0.01), fb)
    fb.close()
    fb = ctx.framebuffer(v.width, v.height)
    with pytest.raises(pcr.PcrError, match="edl_strength"):
        ctx.resolve(cloud, pcr.params(pcr.MODE_LAS2, edl_strength=-1.0), fb)
    fb.close()
