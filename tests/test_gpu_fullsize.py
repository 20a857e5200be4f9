"""GPU parity at a full configuration size: the 100M-point adaptive-precision
terrain (2 km x 2 km, 6-octave fBm, 150 m amplitude, Morton-sorted 10,240-point
batches) rendered at 1920x1080 along a seeded flythrough 100 m above the
ground, compared bit for bit with the multithreaded oracle."""
import math
import os

import numpy as np
import pytest

import compute_rasterizer_amd as pcr
from oracle import orc

pytestmark = pytest.mark.gpu


def flythrough_100m(n, seed=64, extent=2000.0, amplitude=150.0):
    rng = np.random.default_rng(seed)
    lib = orc.load()
    views = []
    for k in range(n):
        a = 2 * math.pi * k / n
        x = extent / 2 + 600 * math.cos(a) + rng.uniform(-20, 20)
        y = extent / 2 + 600 * math.sin(a) + rng.uniform(-20, 20)
        z = lib.orc_terrain_height(100, extent, amplitude, x, y) + 100.0
        ahead = a + math.pi / 2
        tgt = (x + 300 * math.cos(ahead), y + 300 * math.sin(ahead), z - 60.0)
        views.append(pcr.look_at((x, y, z), tgt, up=(0, 0, 1), near=1.0, far=5000.0, width=1920, height=1080))
    return views


@pytest.fixture(scope="module")
def terrain100m(ctx):
    cfg = pcr.gen_cfg(pcr.SCENE_TERRAIN, 100_000_000, seed=100)
    cloud = ctx.generate(cfg)
    lay = cloud.download()
    yield cloud, lay
    cloud.close()


@pytest.mark.parametrize("mode,tile", [(pcr.MODE_LAS2, 2048), (pcr.MODE_LAS, 0)])
def test_config2_flythrough_bit_exact(ctx, terrain100m, mode, tile):
    cloud, lay = terrain100m
    assert cloud.info().num_batches == 9766
    views = flythrough_100m(64)
    fb = ctx.framebuffer(1920, 1080)
    prm = pcr.params(mode, tile_px=tile)
    threads = min(16, len(os.sched_getaffinity(0)))
    for k in (0, 21, 42, 63):
        ctx.render(cloud, views[k], prm, fb)
        ctx.resolve(cloud, prm, fb)
        ref = orc.render_batched_mt(lay["batches"], lay["lo"], lay["med"], lay["hi"], views[k], 1, threads)
        keys = fb.keys()
        assert np.array_equal(keys, ref), f"frame {k}: {(keys != ref).sum()} keys differ"
        assert np.array_equal(fb.rgba(), orc.resolve_index(ref, lay["rgba"], prm.background))
        assert (ref != pcr.FB_CLEAR).sum() > 100_000
