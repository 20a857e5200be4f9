"""GPU: LOD + hqs (PCR_MODE_NODES_HQS, compute_loop_nodes_hqs) equals the
oracle's depth and accumulators bit for bit, and the window-weighted resolve
equals the oracle's float resolve for windows 0..2."""
import numpy as np
import pytest

import compute_rasterizer_amd as pcr
from oracle import orc
from compute_rasterizer_amd.datasets import make_potree

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def octree(ctx):
    xyz, rgba = orc.generate(pcr.gen_cfg(pcr.SCENE_TERRAIN, 300_000, seed=31))
    cloud = ctx.octree_host(xyz, rgba, seed=5, max_node_points=5000)
    lay = cloud.download()
    yield cloud, lay
    cloud.close()


@pytest.mark.parametrize("rule,window,tile", [(pcr.LOD_SPACING, 0, 0), (pcr.LOD_SPACING, 1, 0), (pcr.LOD_SIZE80, 2, 0),
                                              (pcr.LOD_SPACING, 1, 2048), (pcr.LOD_SIZE80, 2, 1024)])
def test_nodes_hqs_matches_oracle(ctx, octree, rule, window, tile):
    cloud, lay = octree
    v = pcr.look_at((1000, -300, 500), (1000, 1000, 0), up=(0, 0, 1), near=1, far=5000, width=640, height=480)
    fb = ctx.framebuffer(v.width, v.height, pcr.FB_HQS)
    prm = pcr.params(pcr.MODE_NODES_HQS, lod_rule=rule, lod_px=1.0, resolve_window=window, tile_px=tile)
    ctx.render(cloud, v, prm, fb)
    ctx.resolve(cloud, prm, fb)
    depth, acc = fb.hqs()
    rdepth, racc = orc.hqs_nodes(lay["batches"], lay["lo"], lay["med"], lay["hi"], lay["rgba"], v, 1, rule, 1.0)
    assert (racc[3::4] > 0).sum() > 1000
    assert np.array_equal(depth, rdepth) and np.array_equal(acc, racc)
    assert np.array_equal(fb.rgba(), orc.hqs_resolve_window(rdepth, racc, v.width, v.height, window))
    fb.close()


def test_nodes_hqs_on_potree_data(ctx):
    meta, hier, oct_, nodes = make_potree(seed=21, max_depth=5)
    cloud = ctx.from_potree(meta, hier, oct_)
    lay = cloud.download()
    info = pcr.potree_info(meta)
    o = tuple(info.min)
    v = pcr.look_at((o[0] + 512, o[1] - 500, o[2] + 700), (o[0] + 512, o[1] + 512, o[2] + 200), up=(0, 0, 1),
                    near=1, far=10000, origin=o, width=800, height=600)
    fb = ctx.framebuffer(v.width, v.height, pcr.FB_HQS)
    prm = pcr.params(pcr.MODE_NODES_HQS, lod_px=1.0, resolve_window=1)
    ctx.render(cloud, v, prm, fb)
    ctx.resolve(cloud, prm, fb)
    depth, acc = fb.hqs()
    rdepth, racc = orc.hqs_nodes(lay["batches"], lay["lo"], lay["med"], lay["hi"], lay["rgba"], v, 1, 0, 1.0)
    assert (racc[3::4] > 0).sum() > 500
    assert np.array_equal(depth, rdepth) and np.array_equal(acc, racc)
    assert np.array_equal(fb.rgba(), orc.hqs_resolve_window(rdepth, racc, v.width, v.height, 1))
    fb.close()
    cloud.close()
