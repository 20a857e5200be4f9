"""GPU: batches that start at any point index. The reference's sparse
buffers put each LAS file's batches at the running point count
(LasLoaderSparse.cpp:524-533), so a batch may start at an index that is not a
multiple of 4. Here every batch is moved to its own offset (0..4 junk words in
front of it, random values), and each render path must still equal the
oracle, which reads exactly [first_point, first_point + num_points)."""
import numpy as np
import pytest

import compute_rasterizer_amd as pcr
from oracle import orc


def shifted(b, lo, med, hi, rgba, seed=3):
    rng = np.random.default_rng(seed)
    nb = len(b)
    out_b = b.copy()
    parts = {k: [] for k in ("lo", "med", "hi", "rgba")}
    src = {"lo": lo, "med": med, "hi": hi, "rgba": rgba}
    pos = 0
    for k in range(nb):
        gap = (k * 7 + 1) % 5   # 0..4 junk words
        f, n = int(b[k]["first_point"]), int(b[k]["num_points"])
        for name in parts:
            parts[name].append(rng.integers(0, 2**32, gap, dtype=np.uint64).astype(np.uint32))
            parts[name].append(src[name][f:f + n])
        out_b[k]["first_point"] = pos + gap
        pos += gap + n
    arrs = {k: np.concatenate(v).astype(np.uint32) for k, v in parts.items()}
    return out_b, arrs["lo"], arrs["med"], arrs["hi"], arrs["rgba"]


@pytest.fixture(scope="module")
def layout():
    xyz, rgba = orc.generate(pcr.gen_cfg(pcr.SCENE_CITY, 120_003, seed=23, num_buildings=12))
    perm = orc.morton_order(xyz)[0]
    xyz, rgba = xyz[perm], rgba[perm]
    b, lo, med, hi = pcr.encode_batches_host(xyz, 2048)
    sb, slo, smed, shi, srgba = shifted(b, lo, med, hi, rgba)
    assert any(int(f) % 4 for f in sb["first_point"])
    return sb, slo, smed, shi, srgba


def cams():
    return [pcr.look_at((150, -120, 160), (150, 150, 10), up=(0, 0, 1), near=1, far=3000, width=640, height=480),
            pcr.look_at((20, 20, 40), (200, 200, 0), up=(0, 0, 1), near=1, far=3000, width=800, height=600)]


def test_oracle_images_do_not_depend_on_offsets(layout):
    """CPU: moving the batches changes the keys' point indices only, so the
    resolved images of the moved and the packed layout are equal."""
    b, lo, med, hi, rgba = layout
    xyz, rgba0 = orc.generate(pcr.gen_cfg(pcr.SCENE_CITY, 120_003, seed=23, num_buildings=12))
    perm = orc.morton_order(xyz)[0]
    b0, lo0, med0, hi0 = pcr.encode_batches_host(xyz[perm], 2048)
    for v in cams():
        ref0, _ = orc.render_batched(b0, lo0, med0, hi0, v, 1)
        ref, _ = orc.render_batched(b, lo, med, hi, v, 1)
        assert np.array_equal(ref0 >> np.uint64(32), ref >> np.uint64(32))
        assert np.array_equal(orc.resolve_index(ref0, rgba0[perm], pcr.BG_LAS),
                              orc.resolve_index(ref, rgba, pcr.BG_LAS))


@pytest.mark.gpu
@pytest.mark.parametrize("mode,tile", [(pcr.MODE_LAS, 0), (pcr.MODE_LAS2, 0), (pcr.MODE_LAS2, 2048)])
def test_unaligned_batches_match_oracle(ctx, layout, mode, tile):
    b, lo, med, hi, rgba = layout
    cloud = ctx.upload_batched(b, lo, med, hi, rgba)
    for v in cams():
        fb = ctx.framebuffer(v.width, v.height)
        prm = pcr.params(mode, tile_px=tile)
        ctx.render(cloud, v, prm, fb)
        ctx.resolve(cloud, prm, fb)
        ref, _ = orc.render_batched(b, lo, med, hi, v, 1)
        assert np.array_equal(fb.keys(), ref)
        assert np.array_equal(fb.rgba(), orc.resolve_index(ref, rgba, prm.background))
        fb.close()
    cloud.close()


@pytest.mark.gpu
@pytest.mark.parametrize("tile", [0, 2048])
def test_unaligned_batches_hqs(ctx, layout, tile):
    b, lo, med, hi, rgba = layout
    cloud = ctx.upload_batched(b, lo, med, hi, rgba)
    v = cams()[0]
    fb = ctx.framebuffer(v.width, v.height, pcr.FB_HQS)
    prm = pcr.params(pcr.MODE_HQS, tile_px=tile)
    ctx.render(cloud, v, prm, fb)
    depth, acc = fb.hqs()
    rdepth, racc = orc.hqs_batched(b, lo, med, hi, rgba, v, 1)
    assert np.array_equal(depth, rdepth) and np.array_equal(acc, racc)
    fb.close()
    cloud.close()


@pytest.mark.gpu
@pytest.mark.parametrize("tune", [(128, 40, 0, 2048), (256, 40, 0, 0)])
def test_unaligned_batches_multiview(ctx, layout, tune):
    b, lo, med, hi, rgba = layout
    cloud = ctx.upload_batched(b, lo, med, hi, rgba)
    views = cams()
    fbs = [ctx.framebuffer(v.width, v.height) for v in views]
    mv = pcr.MultiView(ctx, cloud, views, fbs, pcr.params(pcr.MODE_LAS2), pcr.capi.Tune(*tune), use_graph=False)
    mv.launch()
    ctx.synchronize()
    for v, fb in zip(views, fbs):
        ref, _ = orc.render_batched(b, lo, med, hi, v, 1)
        assert np.array_equal(fb.keys(), ref)
    mv.close()
    for fb in fbs:
        fb.close()
    cloud.close()
