"""GPU: the persistent tiled kernels' work queue at its edges, against the oracle.

k_raster_tiled and k_hqs_tiled take two queue items per atomic, and the last
min(drawn, grid) batches of a frame are drawn as two halves each, all first
halves before the second ones. The persistent grid is min(occupancy x CUs,
batches of the cloud), so a cloud of more batches than the grid exercises
both the whole items and the halves; batch ranges (batch_begin/batch_end)
give the frame 0, 1, 2, 3, 7 and about grid-many drawn batches. Every frame
must equal the oracle's over the same batches, bit for bit.
"""
import numpy as np
import pytest

import compute_rasterizer_amd as pcr
from oracle import orc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def fine_terrain(ctx):
    """2.5M terrain points in 512-point batches: about 4,900 batches, more
    than the persistent grid of the tiled kernels."""
    cfg = pcr.gen_cfg(pcr.SCENE_TERRAIN, 2_500_001, seed=21)
    cloud = ctx.generate(cfg, batch_size=512)
    yield cloud, cloud.download()
    cloud.close()


def overview(w=640, h=360):
    """The whole 2 km terrain in view, so that every batch of a range is drawn."""
    return pcr.look_at((1000, -1800, 2200), (1000, 1000, 0), up=(0, 0, 1), near=1.0, far=10000.0, width=w,
                       height=h)


def ranges(nb):
    return [(5, 5), (7, 8), (100, 102), (2000, 2003), (40, 47), (300, 2347), (300, 2348), (300, 2349),
            (0, nb)]


def test_cloud_has_more_batches_than_the_grid(fine_terrain):
    cloud, lay = fine_terrain
    assert cloud.info().num_batches > 4096


@pytest.mark.parametrize("block", [256, 128])
def test_tiled_raster_over_batch_ranges(ctx, fine_terrain, block):
    cloud, lay = fine_terrain
    nb = cloud.info().num_batches
    view = overview()
    fb = ctx.framebuffer(view.width, view.height)
    for b0, b1 in ranges(nb):
        prm = pcr.params(pcr.MODE_LAS2, tile_px=2048, block_size=block, batch_begin=b0, batch_end=b1)
        ctx.render(cloud, view, prm, fb)
        st = ctx.stats()
        ref, levels = orc.render_batched(lay["batches"][b0:b1], lay["lo"], lay["med"], lay["hi"], view, 1)
        drawn = int((levels != 255).sum())
        assert drawn == b1 - b0, (b0, b1, drawn)   # the overview draws every batch of the range
        assert sum(st.batches_per_level) == drawn
        keys = fb.keys()
        assert np.array_equal(keys, ref), (b0, b1, int((keys != ref).sum()))
    fb.close()


def test_tiled_hqs_over_batch_ranges(ctx, fine_terrain):
    cloud, lay = fine_terrain
    nb = cloud.info().num_batches
    view = overview(480, 270)
    fb = ctx.framebuffer(view.width, view.height, pcr.FB_HQS)
    for b0, b1 in ranges(nb):
        prm = pcr.params(pcr.MODE_HQS, tile_px=2048, batch_begin=b0, batch_end=b1)
        ctx.render(cloud, view, prm, fb)
        d, a = fb.hqs()
        rd, ra = orc.hqs_batched(lay["batches"][b0:b1], lay["lo"], lay["med"], lay["hi"], lay["rgba"], view, 1)
        assert np.array_equal(d, rd), (b0, b1)
        assert np.array_equal(a, ra), (b0, b1)
    fb.close()
