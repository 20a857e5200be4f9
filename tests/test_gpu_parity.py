"""GPU parity: libpcr.so kernels vs the CPU oracle, bit-exact.

Every test builds its input on the device from a seeded generator, reads the
layout back and runs the oracle over the same bytes, so the comparison covers
generation, Morton sort, encoding, cull, precision select, raster and resolve.
"""
import numpy as np
import pytest

import compute_rasterizer_amd as pcr
from oracle import orc

pytestmark = pytest.mark.gpu

C1_EYE, C1_TARGET = (0.5, 0.5, 2.5), (0.5, 0.5, 0.5)


def terrain_views(w=640, h=360, n=4, seed=3):
    """Seeded cameras 100 m above a 2 km terrain, looking ahead and down."""
    rng = np.random.default_rng(seed)
    views = []
    for _ in range(n):
        x, y = rng.uniform(200, 1800, 2)
        ang = rng.uniform(0, 2 * np.pi)
        eye = (x, y, 180.0)
        tgt = (x + 300 * np.cos(ang), y + 300 * np.sin(ang), 60.0)
        views.append(pcr.look_at(eye, tgt, up=(0, 0, 1), near=1.0, far=4000.0, width=w, height=h))
    return views


@pytest.fixture(scope="module")
def terrain(ctx):
    cfg = pcr.gen_cfg(pcr.SCENE_TERRAIN, 1_000_003, seed=11)
    cloud = ctx.generate(cfg)
    yield cloud, cloud.download()
    cloud.close()


@pytest.mark.parametrize("scene,n", [(pcr.SCENE_CUBE, 100_000), (pcr.SCENE_TERRAIN, 100_001),
                                     (pcr.SCENE_CITY, 120_000)])
def test_generator_matches_oracle(ctx, scene, n):
    cfg = pcr.gen_cfg(scene, n, seed=5, num_buildings=16)
    cloud = ctx.generate(cfg, morton_sort=False, keep_points16=True)
    got = cloud.download(batched=False, points16=True)
    xyz, rgba = orc.generate(cfg)
    assert np.array_equal(got["xyz16"].view(np.uint32), xyz.view(np.uint32))
    assert np.array_equal(got["rgba16"], rgba)
    cloud.close()


def test_morton_build_matches_oracle(ctx):
    cfg = pcr.gen_cfg(pcr.SCENE_TERRAIN, 300_000, seed=2)
    cloud = ctx.generate(cfg)
    lay = cloud.download()
    xyz, rgba = orc.generate(cfg)
    perm, keys = orc.morton_order(xyz)
    b, lo, med, hi = orc.encode_batches(xyz[perm], 10240)
    assert np.array_equal(lay["batches"].view(np.uint8), b.view(np.uint8))
    for got, ref in ((lay["lo"], lo), (lay["med"], med), (lay["hi"], hi), (lay["rgba"], rgba[perm])):
        assert np.array_equal(got, ref)


def test_plain16_config1(ctx):
    """1M points U[0,1)^3, colour payload, 1280x720 (compute_2021_earlyz)."""
    cfg = pcr.gen_cfg(pcr.SCENE_CUBE, 1_000_000, seed=1)
    xyz, rgba = orc.generate(cfg)
    cloud = ctx.upload_points16(xyz, rgba)
    view = pcr.look_at(C1_EYE, C1_TARGET, width=1280, height=720)
    prm = pcr.params(pcr.MODE_PLAIN16)
    fb = ctx.framebuffer(1280, 720)
    ctx.render(cloud, view, prm, fb)
    ctx.resolve(cloud, prm, fb)
    ref = orc.render_points16(xyz, rgba, view)
    assert np.array_equal(fb.keys(), ref)
    assert np.array_equal(fb.rgba(), orc.resolve_color(ref, pcr.BG_BLACK))
    assert (ref != pcr.FB_CLEAR).sum() > 100_000


@pytest.mark.parametrize("mode,tile", [(pcr.MODE_LAS, 0), (pcr.MODE_LAS2, 0), (pcr.MODE_LAS2, 2048)])
@pytest.mark.parametrize("culling", [1, 0])
def test_batched_matches_oracle(ctx, terrain, mode, tile, culling):
    cloud, lay = terrain
    fb = ctx.framebuffer(640, 360)
    prm = pcr.params(mode, culling=culling, tile_px=tile)
    for view in terrain_views():
        ctx.render(cloud, view, prm, fb)
        ctx.resolve(cloud, prm, fb)
        st = ctx.stats()
        ref, levels = orc.render_batched(lay["batches"], lay["lo"], lay["med"], lay["hi"], view, culling)
        keys = fb.keys()
        assert np.array_equal(keys, ref), f"{(keys != ref).sum()} keys differ"
        assert np.array_equal(fb.rgba(), orc.resolve_index(ref, lay["rgba"], prm.background))
        hist = [int((levels == k).sum()) for k in range(5)]
        assert list(st.batches_per_level) == hist
        assert st.batches_culled == int((levels == 255).sum())


def test_precision_levels_all_used(ctx, terrain):
    """A low camera close to the ground sees batches at several precision levels."""
    cloud, lay = terrain
    view = pcr.look_at((1000, 1000, 100), (1000, 1010, 60), up=(0, 0, 1), near=0.5, far=4000, width=1920,
                       height=1080)
    fb = ctx.framebuffer(1920, 1080)
    prm = pcr.params(pcr.MODE_LAS2)
    ctx.render(cloud, view, prm, fb)
    ref, levels = orc.render_batched(lay["batches"], lay["lo"], lay["med"], lay["hi"], view, 1)
    assert np.array_equal(fb.keys(), ref)
    assert len(set(levels.tolist()) - {255}) >= 2


@pytest.mark.parametrize("tile,block", [(0, 0), (2048, 256), (1024, 128), (4096, 256)])
def test_hqs_batched(ctx, terrain, tile, block):
    """Direct and LDS-tiled hqs passes against the oracle."""
    cloud, lay = terrain
    fb = ctx.framebuffer(640, 360, pcr.FB_HQS)
    prm = pcr.params(pcr.MODE_HQS, tile_px=tile, block_size=block)
    for view in terrain_views(n=2, seed=9):
        ctx.render(cloud, view, prm, fb)
        ctx.resolve(cloud, prm, fb)
        d, a = fb.hqs()
        rd, ra = orc.hqs_batched(lay["batches"], lay["lo"], lay["med"], lay["hi"], lay["rgba"], view, 1)
        assert np.array_equal(d, rd)
        assert np.array_equal(a, ra)
        assert np.array_equal(fb.rgba(), orc.hqs_resolve(ra, prm.background))


@pytest.mark.parametrize("tile", [0, 2048])
def test_hqs_dense_far_view(ctx, terrain, tile):
    """The whole terrain from far away at 160x90: each pixel gathers
    hundreds of points within 1.01x of its depth, so the tiled pass 2 merges
    many colours per pixel in LDS before its global adds."""
    cloud, lay = terrain
    view = pcr.look_at((1000, -6000, 3000), (1000, 1000, 0), up=(0, 0, 1), near=1.0, far=20000.0, width=160,
                       height=90)
    fb = ctx.framebuffer(160, 90, pcr.FB_HQS)
    prm = pcr.params(pcr.MODE_HQS, tile_px=tile)
    ctx.render(cloud, view, prm, fb)
    ctx.resolve(cloud, prm, fb)
    d, a = fb.hqs()
    rd, ra = orc.hqs_batched(lay["batches"], lay["lo"], lay["med"], lay["hi"], lay["rgba"], view, 1)
    assert ra[3::4].max() > 200
    assert np.array_equal(d, rd) and np.array_equal(a, ra)
    assert np.array_equal(fb.rgba(), orc.hqs_resolve(ra, prm.background))
    fb.close()


def test_hqs16_config3(ctx):
    """hqs on config-1 data at 1920x1080."""
    cfg = pcr.gen_cfg(pcr.SCENE_CUBE, 1_000_000, seed=1)
    xyz, rgba = orc.generate(cfg)
    cloud = ctx.upload_points16(xyz, rgba)
    view = pcr.look_at(C1_EYE, C1_TARGET, width=1920, height=1080)
    prm = pcr.params(pcr.MODE_HQS16)
    fb = ctx.framebuffer(1920, 1080, pcr.FB_HQS)
    ctx.render(cloud, view, prm, fb)
    ctx.resolve(cloud, prm, fb)
    d, a = fb.hqs()
    rd, ra = orc.hqs_points16(xyz, rgba, view)
    assert np.array_equal(d, rd)
    assert np.array_equal(a, ra)
    assert np.array_equal(fb.rgba(), orc.hqs_resolve(ra, prm.background))
    assert a[3::4].sum() >= (rd != 0xFFFFFFFF).sum()


def test_host_encoded_upload(ctx):
    """The drop-in path: a layout encoded on the host (LasLoaderSparse's job) uploaded as is."""
    cfg = pcr.gen_cfg(pcr.SCENE_TERRAIN, 50_001, seed=4)
    xyz, rgba = orc.generate(cfg)
    b, lo, med, hi = pcr.encode_batches_host(xyz, 10240)
    rb, rlo, rmed, rhi = orc.encode_batches(xyz, 10240)
    assert np.array_equal(b.view(np.uint8), rb.view(np.uint8))
    assert np.array_equal(lo, rlo) and np.array_equal(med, rmed) and np.array_equal(hi, rhi)
    cloud = ctx.upload_batched(b, lo, med, hi, rgba)
    view = pcr.look_at((1000, -300, 500), (1000, 1000, 0), up=(0, 0, 1), near=1, far=5000, width=800, height=600)
    fb = ctx.framebuffer(800, 600)
    for mode in (pcr.MODE_LAS, pcr.MODE_LAS2):
        prm = pcr.params(mode)
        ctx.render(cloud, view, prm, fb)
        ref, _ = orc.render_batched(b, lo, med, hi, view, 1)
        assert np.array_equal(fb.keys(), ref)


def test_empty_frame(ctx, terrain):
    cloud, lay = terrain
    view = pcr.look_at((1000, 1000, 500), (1000, 1000, 1000), up=(0, 1, 0), near=1, far=100, width=320, height=200)
    fb = ctx.framebuffer(320, 200)
    prm = pcr.params(pcr.MODE_LAS2)
    ctx.render(cloud, view, prm, fb)
    ctx.resolve(cloud, prm, fb)
    assert (fb.keys() == pcr.FB_CLEAR).all()
    assert (fb.rgba() == prm.background).all()
    st = ctx.stats()
    assert st.batches_culled == st.batches_total


@pytest.mark.parametrize("tile_px,block", [(1024, 256), (2048, 256), (4096, 256), (2048, 128), (4096, 128)])
def test_tiled_raster_matches_oracle(ctx, terrain, tile_px, block):
    """LDS-tiled merge gives the same bits as the direct path, far and close views."""
    cloud, lay = terrain
    prm = pcr.params(pcr.MODE_LAS2, tile_px=tile_px, block_size=block)
    views = terrain_views(n=3, seed=21) + [
        pcr.look_at((1000, 1000, 100), (1000, 1010, 60), up=(0, 0, 1), near=0.5, far=4000, width=640, height=360)]
    fb = ctx.framebuffer(640, 360)
    for view in views:
        ctx.render(cloud, view, prm, fb)
        ref, _ = orc.render_batched(lay["batches"], lay["lo"], lay["med"], lay["hi"], view, 1)
        keys = fb.keys()
        assert np.array_equal(keys, ref), f"{(keys != ref).sum()} keys differ"


@pytest.mark.parametrize("n,maxpts", [(400_000, 20000), (150_001, 3000)])
def test_device_octree_equals_host_octree(ctx, n, maxpts):
    cfg = pcr.gen_cfg(pcr.SCENE_CITY, n, seed=23, num_buildings=8)
    xyz, rgba = orc.generate(cfg)
    host = pcr.octree_layout_host(xyz, rgba, seed=6, max_node_points=maxpts)
    cloud = ctx.octree_generate(cfg, seed=6, max_node_points=maxpts)
    dev = cloud.download()
    assert np.array_equal(dev["batches"].view(np.uint8), host[0].view(np.uint8))
    for k, h in zip(("lo", "med", "hi", "rgba"), host[1:]):
        assert np.array_equal(dev[k], h), k
    assert cloud.info().num_levels == int(host[0]["level"].max()) + 1
    cloud.close()


@pytest.fixture(scope="module")
def octree(ctx):
    cfg = pcr.gen_cfg(pcr.SCENE_TERRAIN, 400_000, seed=13)
    xyz, rgba = orc.generate(cfg)
    cloud = ctx.octree_host(xyz, rgba, seed=3, max_node_points=20000)
    lay = cloud.download()
    yield cloud, lay
    cloud.close()


@pytest.mark.parametrize("rule", [pcr.LOD_SPACING, pcr.LOD_SIZE80])
@pytest.mark.parametrize("tile_px", [0, 4096])
def test_nodes_lod_matches_oracle(ctx, octree, rule, tile_px):
    cloud, lay = octree
    assert cloud.info().num_levels >= 2
    fb = ctx.framebuffer(960, 540)
    prm = pcr.params(pcr.MODE_NODES, lod_rule=rule, lod_px=1.0, tile_px=tile_px)
    views = terrain_views(w=960, h=540, n=3, seed=31) + [
        pcr.look_at((1000, -3000, 4000), (1000, 1000, 0), up=(0, 0, 1), near=1, far=20000, width=960, height=540)]
    for view in views:
        ctx.render(cloud, view, prm, fb)
        ctx.resolve(cloud, prm, fb)
        ref, flags = orc.render_nodes(lay["batches"], lay["lo"], lay["med"], lay["hi"], view, 1, rule, 1.0)
        assert np.array_equal(ctx.lod_selection(cloud), flags)
        assert np.array_equal(fb.keys(), ref)
        assert np.array_equal(fb.rgba(), orc.resolve_index(ref, lay["rgba"], prm.background))
