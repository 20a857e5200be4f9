"""GPU: LAZ ingest. pcr_cloud_from_las decodes LAZ files on the host (the
library's own LASzip-format decoder) and encodes them on the device the way
LasLoaderSparse::loadLaz does (LasLoaderSparse.cpp:214-330): the layout must
equal the oracle's encoding of the same records from the uncompressed LAS
file, with colour 0 for point formats without R, G, B (loadLaz takes colour
from the LASzip point). The fixtures were written by the reference's vendored
LASzip (tests/golden/make_laz.py)."""
import os

import numpy as np
import pytest

import compute_rasterizer_amd as pcr
from oracle import orc

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = sorted(f[:-4] for f in os.listdir(os.path.join(GOLDEN, "laz")) if f.endswith(".laz"))


@pytest.fixture(scope="module")
def store():
    return np.load(os.path.join(GOLDEN, "laz_records.npz"))


def pair(store, name):
    """(laz bytes, las bytes) of a fixture."""
    key = name[:-3] + "_v3" if name.endswith("_v4") else name
    with open(os.path.join(GOLDEN, "laz", name + ".laz"), "rb") as f:
        laz = f.read()
    las = store[key + "_las_header"].tobytes() + store[key + "_records"].tobytes()
    return laz, las


def oracle_files(datas, lazflags):
    parts, sparse = [], 0
    for k, (d, z) in enumerate(zip(datas, lazflags)):
        info = orc.las_parse_header(d)
        parts.append(orc.las_encode(d, info, sparse_offset=sparse, file_index=k, laz=z))
        sparse += int(info.num_points)
    return tuple(np.concatenate([p[i] for p in parts]) for i in range(5))


def check_layout(cloud, ref):
    lay = cloud.download()
    b, lo, med, hi, rgba = ref
    assert np.array_equal(lay["batches"].view(np.uint8), b.view(np.uint8))
    for k, r in (("lo", lo), ("med", med), ("hi", hi), ("rgba", rgba)):
        assert np.array_equal(lay[k], r), k


@pytest.mark.parametrize("name", CASES)
def test_laz_file_encodes_like_the_oracle(ctx, store, name):
    laz, las = pair(store, name)
    print(f"case {name}", flush=True)
    cloud, first = ctx.from_las([laz])
    check_layout(cloud, oracle_files([las], [True]))
    assert int(first[1]) == (int(orc.las_parse_header(las).num_points) + 10239) // 10240
    cloud.close()


def test_mixed_las_and_laz_files(ctx, store):
    names = ["f2_rgb_chunk1500", "f7_v3", "f0_chunk1000", "f8_extra2_v4", "f3_extra3_variable"]
    pairs = [pair(store, n) for n in names]
    datas = [pairs[0][1]] + [p[0] for p in pairs[1:]]      # the first file uncompressed
    cloud, first = ctx.from_las(datas)
    check_layout(cloud, oracle_files([p[1] for p in pairs], [False] + [True] * (len(pairs) - 1)))
    cloud.close()


def test_laz_cloud_renders_like_its_las_cloud(ctx, store):
    laz, las = pair(store, "f7_v4")
    a, _ = ctx.from_las([laz])
    b, _ = ctx.from_las([las])
    info = orc.las_parse_header(las)
    rec = store["f7_v3_records"]
    xyz = rec[:, :12].copy().view(np.int32).reshape(-1, 3) * np.array(info.scale) + np.array(info.offset)
    c = np.median(xyz, axis=0) - np.array(info.min)      # positions are relative to the header min
    v = pcr.look_at((c[0], c[1] - 30.0, c[2] + 40.0), tuple(c), up=(0, 0, 1), near=0.5, far=5000.0, width=320,
                    height=240)
    out = []
    for cloud in (a, b):
        fb = ctx.framebuffer(320, 240)
        ctx.render(cloud, v, pcr.params(pcr.MODE_LAS2), fb)
        out.append(fb.keys())
        fb.close()
        cloud.close()
    assert (out[0] != pcr.FB_CLEAR).sum() > 100
    assert np.array_equal(out[0], out[1])


def test_multi_chunk_laz_files_on_device(ctx, store):
    """Larger LAZ files (hundreds of chunks, decoded on several host threads in
    pinned slabs) encode like the oracle's encoding of the same LAS records."""
    from compute_rasterizer_amd.datasets import laz_tile
    datas, lases = [], []
    for name, key, copies in (("f7_v4", "f7_v3", 60), ("f3_extra3_variable", "f3_extra3_variable", 41)):
        laz, _ = pair(store, name)
        datas.append(laz_tile(laz, copies))
        rec = store[key + "_records"]
        hdr = bytearray(store[key + "_las_header"].tobytes())
        n = copies * rec.shape[0]
        hdr[107:111] = np.uint32(n).tobytes()
        if len(hdr) >= 255:
            hdr[247:255] = np.uint64(n).tobytes()
        lases.append(bytes(hdr) + np.tile(rec, (copies, 1)).tobytes())
    print("points", [int(pcr.laz_info(d).num_points) for d in datas], flush=True)
    cloud, first = ctx.from_las(datas)
    check_layout(cloud, oracle_files(lases, [True, True]))
    cloud.close()
