"""GPU: the Potree 2.0 reader builds on the device the oracle's batches and
planes byte for byte, and LOD frames of the result equal the oracle's
(compute_loop_nodes/render.cs:244-479 with both LOD rules)."""
import numpy as np
import pytest

import compute_rasterizer_amd as pcr
from oracle import orc
from compute_rasterizer_amd.datasets import make_potree

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dataset():
    return make_potree(seed=11, max_depth=5, chunk_depth=2)


def test_device_equals_oracle(ctx, dataset):
    meta, hier, octree, nodes = dataset
    cloud = ctx.from_potree(meta, hier, octree)
    info = pcr.potree_info(meta)
    b, boxes = orc.potree_nodes(info, hier)
    lo, med, hi, rgba = orc.potree_encode(info, octree, b, boxes)
    lay = cloud.download()
    assert np.array_equal(lay["batches"].view(np.uint8), b.view(np.uint8))
    for k, ref in (("lo", lo), ("med", med), ("hi", hi), ("rgba", rgba)):
        assert np.array_equal(lay[k], ref), k
    cloud.close()


@pytest.mark.parametrize("rule,tile", [(pcr.LOD_SPACING, 0), (pcr.LOD_SPACING, 2048), (pcr.LOD_SIZE80, 2048)])
def test_lod_frames_equal_oracle(ctx, dataset, rule, tile):
    meta, hier, octree, nodes = dataset
    cloud = ctx.from_potree(meta, hier, octree)
    info = pcr.potree_info(meta)
    b, boxes = orc.potree_nodes(info, hier)
    lo, med, hi, rgba = orc.potree_encode(info, octree, b, boxes)
    origin = tuple(info.min)
    for eye, tgt in [((512.0, -600.0, 700.0), (512.0, 512.0, 300.0)), ((100.0, 100.0, 900.0), (600.0, 700.0, 0.0))]:
        world = lambda p: tuple(p[a] + origin[a] for a in range(3))
        v = pcr.look_at(world(eye), world(tgt), up=(0, 0, 1), near=1.0, far=10000.0, origin=origin, width=960,
                        height=540)
        prm = pcr.params(pcr.MODE_NODES, lod_rule=rule, lod_px=1.0, tile_px=tile)
        fb = ctx.framebuffer(v.width, v.height)
        ctx.render(cloud, v, prm, fb)
        ctx.resolve(cloud, prm, fb)
        ref, _ = orc.render_nodes(b, lo, med, hi, v, 1, rule, 1.0)
        assert (ref != pcr.FB_CLEAR).sum() > 1000
        assert np.array_equal(fb.keys(), ref)
        assert np.array_equal(fb.rgba(), orc.resolve_index(ref, rgba, prm.background))
        fb.close()
    cloud.close()


def test_malformed_hierarchy_is_refused(ctx, dataset):
    meta, hier, octree, nodes = dataset
    with pytest.raises(pcr.PcrError) as e:
        ctx.from_potree(meta, hier[:-22], octree)
    assert "hierarchy" in str(e.value)
