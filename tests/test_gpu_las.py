"""GPU: LAS ingest on the device equals the oracle's LasLoaderSparse::loadLas
restatement byte for byte, and a multi-file cloud renders, file by file
with each file's own origin (compute_loop_las.h:141-199), into the same
framebuffer as the oracle does."""
import ctypes as C
import os

import numpy as np
import pytest

import compute_rasterizer_amd as pcr
from compute_rasterizer_amd import capi
from oracle import orc
from compute_rasterizer_amd.datasets import make_las

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "5points_14.las")


def from_las(ctx, datas):
    bufs = [np.frombuffer(d, np.uint8) for d in datas]
    files = (capi.LasFile * len(bufs))(*[capi.LasFile(b.ctypes.data, b.size) for b in bufs])
    first = np.zeros(len(bufs) + 1, np.uint32)
    h = C.c_void_p()
    capi.check(capi.load().pcr_cloud_from_las(ctx.h, files, len(bufs), capi.ptr(first), C.byref(h)))
    return pcr.Cloud(ctx, h), first


def oracle_las(datas):
    parts, sparse = [], 0
    for k, d in enumerate(datas):
        info = orc.las_parse_header(d)
        parts.append(orc.las_encode(d, info, sparse_offset=sparse, file_index=k))
        sparse += int(info.num_points)
    return tuple(np.concatenate([p[i] for p in parts]) for i in range(5))


@pytest.fixture(scope="module")
def files():
    # odd point counts, so later files start at indices that are not multiples of 4
    return [make_las(25_601, point_format=2, version=(1, 2), seed=1)[0],
            make_las(13_333, point_format=3, version=(1, 3), seed=2, offset=(500100.0, 4000050.0, 90.0))[0],
            make_las(20_482, point_format=7, version=(1, 4), seed=3, rgb16=False)[0],
            make_las(9_999, point_format=8, version=(1, 4), seed=4, flat_z=True, extra_bytes=3)[0]]


def test_device_encode_equals_oracle(ctx, files):
    cloud, first = from_las(ctx, files)
    b, lo, med, hi, rgba = oracle_las(files)
    lay = cloud.download()
    assert np.array_equal(lay["batches"].view(np.uint8), b.view(np.uint8))
    for k, ref in (("lo", lo), ("med", med), ("hi", hi), ("rgba", rgba)):
        assert np.array_equal(lay[k], ref), k
    sizes = [int(orc.las_parse_header(d).num_points) for d in files]
    assert [int(x) for x in first] == list(np.cumsum([0] + [(n + 10239) // 10240 for n in sizes]))
    cloud.close()


def test_example_file_on_device(ctx):
    data = open(GOLDEN, "rb").read()
    cloud, first = from_las(ctx, [data])
    b, lo, med, hi, rgba = oracle_las([data])
    lay = cloud.download()
    assert np.array_equal(lay["batches"].view(np.uint8), b.view(np.uint8))
    assert np.array_equal(lay["lo"], lo) and np.array_equal(lay["rgba"], rgba)
    cloud.close()


@pytest.mark.parametrize("mode,tile", [(pcr.MODE_LAS, 0), (pcr.MODE_LAS2, 2048), (pcr.MODE_LAS2, 0)])
def test_multi_file_render(ctx, files, mode, tile):
    cloud, first = from_las(ctx, files)
    b, lo, med, hi, rgba = oracle_las(files)
    infos = [orc.las_parse_header(d) for d in files]
    w, h = 800, 600
    fb = ctx.framebuffer(w, h)
    ref = None
    base = infos[0]
    for k, info in enumerate(infos):
        # one camera in world space; each file's transform ends with a
        # translation to its own header min (compute_loop_las.h:170-176)
        eye = (base.min[0] + 150.0, base.min[1] - 250.0, base.min[2] + 200.0)
        tgt = (base.min[0] + 150.0, base.min[1] + 100.0, base.min[2])
        v = pcr.look_at(eye, tgt, up=(0, 0, 1), near=1.0, far=5000.0, origin=tuple(info.min), width=w, height=h)
        prm = pcr.params(mode, tile_px=tile, clear_before=int(k == 0), batch_begin=int(first[k]),
                         batch_end=int(first[k + 1]))
        ctx.render(cloud, v, prm, fb)
        ref, _ = orc.render_batched(b[first[k]:first[k + 1]], lo, med, hi, v, 1, fb=ref)
    prm = pcr.params(mode)
    ctx.resolve(cloud, prm, fb)
    keys = fb.keys()
    assert (keys != pcr.FB_CLEAR).sum() > 1000
    assert np.array_equal(keys, ref)
    assert np.array_equal(fb.rgba(), orc.resolve_index(ref, rgba, prm.background))
    fb.close()
    cloud.close()


def test_bad_batch_range_is_refused(ctx, files):
    cloud, first = from_las(ctx, files[:1])
    v = pcr.look_at((0, -100, 100), (0, 0, 0), up=(0, 0, 1), width=64, height=64)
    fb = ctx.framebuffer(64, 64)
    with pytest.raises(pcr.PcrError) as e:
        ctx.render(cloud, v, pcr.params(pcr.MODE_LAS2, batch_begin=1, batch_end=int(first[1]) + 1), fb)
    assert "batch range" in str(e.value)
    fb.close()
    cloud.close()
