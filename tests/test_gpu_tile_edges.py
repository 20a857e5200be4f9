"""GPU: the tiled raster and hqs passes at the edges of their fast paths.

Batches whose screen rectangle keeps a pixel from every image edge project
without the per-point range, NDC and image tests; others keep them. Images of
1 to 5 pixels on a side make one- and two-column (and -row) rectangles, which
take the direct path (one column) or the tile's row split (two and more);
small odd sizes put many rectangles against the edges, and a camera a few
decimetres above the ground puts batches right in front of the eye. Every
framebuffer must equal the oracle's bit for bit.
"""
import numpy as np
import pytest

import compute_rasterizer_amd as pcr
from oracle import orc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def city(ctx):
    cloud = ctx.generate(pcr.gen_cfg(pcr.SCENE_CITY, 400_009, seed=37, num_buildings=24))
    yield cloud, cloud.download()
    cloud.close()


def edge_views():
    out = []
    eye, tgt = (400.0, 250.0, 260.0), (1000.0, 1000.0, 30.0)
    for w, h in [(1, 200), (2, 300), (3, 97), (200, 1), (300, 2), (5, 5), (37, 23), (641, 359)]:
        out.append(pcr.look_at(eye, tgt, up=(0, 0, 1), near=1.0, far=5000.0, width=w, height=h))
    # a camera 0.3 m above the ground between buildings, near plane 1 cm
    out.append(pcr.look_at((1000.0, 1000.0, 0.3), (1040.0, 1010.0, 0.0), up=(0, 0, 1), near=0.01, far=5000.0,
                           width=480, height=270))
    # far view: most batches are a few pixels wide
    out.append(pcr.look_at((1000.0, -5000.0, 2500.0), (1000.0, 1000.0, 0.0), up=(0, 0, 1), near=1.0, far=20000.0,
                           width=97, height=61))
    return out


def test_views_cover_the_edges(city):
    """The oracle draws points in every view, so no case is vacuous."""
    _, lay = city
    drawn = []
    for v in edge_views():
        keys, _ = orc.render_batched(lay["batches"], lay["lo"], lay["med"], lay["hi"], v, 1)
        drawn.append(int((keys != pcr.FB_CLEAR).sum()))
    assert all(d > 0 for d in drawn), drawn


@pytest.mark.parametrize("tile_px,block", [(2048, 256), (1024, 128), (4096, 256)])
def test_tiled_raster_at_edges(ctx, city, tile_px, block):
    cloud, lay = city
    prm = pcr.params(pcr.MODE_LAS2, tile_px=tile_px, block_size=block)
    for v in edge_views():
        fb = ctx.framebuffer(v.width, v.height)
        ctx.render(cloud, v, prm, fb)
        ref, _ = orc.render_batched(lay["batches"], lay["lo"], lay["med"], lay["hi"], v, 1)
        keys = fb.keys()
        assert np.array_equal(keys, ref), f"{v.width}x{v.height}: {(keys != ref).sum()} keys differ"
        fb.close()


@pytest.mark.parametrize("tile_px", [2048, 4096])
def test_tiled_hqs_at_edges(ctx, city, tile_px):
    cloud, lay = city
    prm = pcr.params(pcr.MODE_HQS, tile_px=tile_px)
    for v in edge_views():
        fb = ctx.framebuffer(v.width, v.height, pcr.FB_HQS)
        ctx.render(cloud, v, prm, fb)
        depth, acc = fb.hqs()
        rdepth, racc = orc.hqs_batched(lay["batches"], lay["lo"], lay["med"], lay["hi"], lay["rgba"], v, 1)
        assert np.array_equal(depth, rdepth), f"{v.width}x{v.height}: depth differs"
        assert np.array_equal(acc, racc), f"{v.width}x{v.height}: accumulators differ"
        fb.close()


def test_multiview_tiled_at_edges(ctx, city):
    cloud, lay = city
    vs = edge_views()
    fbs = [ctx.framebuffer(v.width, v.height) for v in vs]
    mv = pcr.MultiView(ctx, cloud, vs, fbs, pcr.params(pcr.MODE_LAS2), pcr.capi.Tune(128, 40, 0, 2048),
                       use_graph=False)
    mv.launch()
    ctx.synchronize()
    for v, fb in zip(vs, fbs):
        ref, _ = orc.render_batched(lay["batches"], lay["lo"], lay["med"], lay["hi"], v, 1)
        assert np.array_equal(fb.keys(), ref), f"{v.width}x{v.height}"
    mv.close()
    for fb in fbs:
        fb.close()
