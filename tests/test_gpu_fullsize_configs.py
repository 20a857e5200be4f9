"""GPU parity at the stated sizes of the hqs, batched multi-view and serving
configurations, against the oracle:

  hqs      1M points xyz ~ U[0,1)^3 (16 B/pt, unsorted), 1920x1080, both passes
           and the integer resolve (depth, four sums per pixel, image);
  batched  the 1e9-point city (256 buildings on a 2 km fBm terrain), 48 seeded
           views of uneven resolution from {256^2, 640x480, 1280x720,
           1920x1080, 2560x1440} rendered by one autotuned hipGraph launch;
           every view's keys and image against the oracle;
  serving  four device-resident LOD clouds (10M, 100M, 500M and 1e9 city
           points, octrees of at most 20k points per node), a seeded request
           trace (random cloud, camera and resolution, 1-px LOD threshold)
           served with a 16 GB budget; every image against the oracle, alone
           and in the mix.
"""
import math
import os

import numpy as np
import pytest

import compute_rasterizer_amd as pcr
from oracle import orc

pytestmark = pytest.mark.gpu

SIZES = [(256, 256), (640, 480), (1280, 720), (1920, 1080), (2560, 1440)]
THREADS = min(16, len(os.sched_getaffinity(0)))


def seeded_views(n, seed, extent=2000.0):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        w, h = SIZES[int(rng.integers(0, len(SIZES)))]
        x, y = rng.uniform(0.15 * extent, 0.85 * extent, 2)
        a = rng.uniform(0, 2 * math.pi)
        eye = (x, y, rng.uniform(150.0, 400.0))
        tgt = (x + 500 * math.cos(a), y + 500 * math.sin(a), rng.uniform(0.0, 80.0))
        out.append(pcr.look_at(eye, tgt, up=(0, 0, 1), fovy=60.0, near=1.0, far=5000.0, width=w, height=h))
    return out


def test_hqs_cube_1m_1080p_equals_oracle(ctx):
    cfg = pcr.gen_cfg(pcr.SCENE_CUBE, 1_000_000, seed=1)
    cloud = ctx.generate(cfg, morton_sort=False, keep_points16=True)
    lay = cloud.download(batched=False, points16=True)
    xyz, rgba = orc.generate(cfg)
    assert np.array_equal(lay["xyz16"], xyz) and np.array_equal(lay["rgba16"], rgba)   # unsorted C1 data
    v = pcr.look_at((0.5, 0.5, 2.5), (0.5, 0.5, 0.5), up=(0, 1, 0), fovy=60.0, near=0.1, far=1000.0, width=1920,
                    height=1080)
    fb = ctx.framebuffer(1920, 1080, pcr.FB_HQS)
    prm = pcr.params(pcr.MODE_HQS16)
    ctx.render(cloud, v, prm, fb)
    ctx.resolve(cloud, prm, fb)
    d, a = fb.hqs()
    rd, ra = orc.hqs_points16(xyz, rgba, v)
    assert np.array_equal(d, rd) and np.array_equal(a, ra)
    assert np.array_equal(fb.rgba(), orc.hqs_resolve(ra, prm.background))
    assert (ra[3::4] > 1).sum() > 1000          # pixels averaging several points
    fb.close()
    cloud.close()


def test_batched_48_views_1e9_points_one_graph_launch(ctx):
    cloud = ctx.generate(pcr.gen_cfg(pcr.SCENE_CITY, 1_000_000_000, seed=42, num_buildings=256, terrain_frac=0.2))
    assert cloud.info().num_batches == 97657
    views = seeded_views(48, seed=7)
    fbs = [ctx.framebuffer(v.width, v.height) for v in views]
    prm = pcr.params(pcr.MODE_LAS2)
    best, _ = pcr.autotune(ctx, cloud, views, fbs, prm, iters=2)
    mv = pcr.MultiView(ctx, cloud, views, fbs, prm, tune=best)
    mv.launch()
    ctx.synchronize()
    print(f"batched: tune {best.block_size}/{best.points_per_thread}/{best.tile_px}", flush=True)
    lay = cloud.download()
    for k, v in enumerate(views):
        ref = orc.render_batched_mt(lay["batches"], lay["lo"], lay["med"], lay["hi"], v, 1, THREADS)
        keys = fbs[k].keys()
        assert np.array_equal(keys, ref), f"view {k} ({v.width}x{v.height}): {(keys != ref).sum()} keys differ"
        assert np.array_equal(fbs[k].rgba(), orc.resolve_index(ref, lay["rgba"], prm.background)), k
    assert sum((fb.keys() != pcr.FB_CLEAR).sum() for fb in fbs) > 10_000_000
    mv.close()
    for fb in fbs:
        fb.close()
    del lay
    cloud.close()


def test_serving_four_lod_clouds_equal_oracle_alone_and_in_mix(ctx):
    sizes = [10_000_000, 100_000_000, 500_000_000, 1_000_000_000]
    clouds, lays = [], []
    for k, n in enumerate(sizes):
        cfg = pcr.gen_cfg(pcr.SCENE_CITY, n, seed=200 + k, num_buildings=256, terrain_frac=0.2)
        clouds.append(ctx.octree_generate(cfg, seed=k + 1, max_node_points=20000))
        lays.append(clouds[-1].download())
        assert lays[-1]["batches"]["num_points"].max() <= 20000
    srv = pcr.Server(ctx, budget_bytes=16 << 30, max_batch=16, num_streams=4)
    ids = [srv.add_cloud(c) for c in clouds]
    rng = np.random.default_rng(13)
    trace = [(int(rng.integers(0, 4)), v) for v in seeded_views(24, seed=12)]
    alone = [srv.wait(srv.submit(ids[c], v, pcr.MODE_NODES, lod_px=1.0))[0] for c, v in trace]
    tickets = [srv.submit(ids[c], v, pcr.MODE_NODES, lod_px=1.0) for c, v in trace]   # all in flight together
    mixed = [srv.wait(t)[0] for t in tickets]
    st = srv.stats()
    assert st.completed == 2 * len(trace) and st.fb_bytes_in_use == 0
    assert st.launches < 2 * len(trace)                  # the mix was batched
    assert st.points_drawn > 0 and st.bytes_algorithmic > 0
    for k, (c, v) in enumerate(trace):
        L = lays[c]
        ref = orc.render_nodes_mt(L["batches"], L["lo"], L["med"], L["hi"], v, 1, pcr.LOD_SPACING, 1.0, THREADS)
        img = orc.resolve_index(ref, L["rgba"], pcr.BG_LAS)
        assert np.array_equal(alone[k], img), f"request {k} (cloud {c}) alone differs from the oracle"
        assert np.array_equal(mixed[k], img), f"request {k} (cloud {c}) in the mix differs from the oracle"
    srv.close()
    del lays
    for c in clouds:
        c.close()
