"""CPU: the oracle's Potree 2.0 reader (PotreeData.h:144-317, 439-511) against
synthetic datasets whose node tree is known: every node's box, point count,
first point and depth, the octree.bin order, the parent links, and the
encoding's round trip. Malformed hierarchies are refused."""
import json

import numpy as np
import pytest

import compute_rasterizer_amd as pcr
from oracle import orc
from compute_rasterizer_amd.datasets import make_potree


@pytest.mark.parametrize("seed,chunk_depth,max_depth", [(1, 2, 4), (2, 1, 3), (3, 3, 5), (4, 8, 3)])
def test_hierarchy_matches_the_written_tree(seed, chunk_depth, max_depth):
    meta, hier, octree, nodes = make_potree(seed=seed, chunk_depth=chunk_depth, max_depth=max_depth)
    info = pcr.potree_info(meta)
    b, boxes = orc.potree_nodes(info, hier)
    assert len(b) == len(nodes)
    # octree.bin order
    assert (np.diff(b["first_point"].astype(np.int64)) > 0).all()
    by_first = {n["first"]: n for n in nodes}
    for k in range(len(b)):
        n = by_first[int(b[k]["first_point"])]
        assert tuple(boxes[k]) == tuple(n["box"])            # power-of-two boxes are exact
        assert int(b[k]["num_points"]) == n["count"] and int(b[k]["level"]) == n["level"]
        assert np.array_equal(b[k]["min"], np.float32(n["box"][:3])) and np.array_equal(b[k]["max"], np.float32(n["box"][3:]))
        p = int(b[k]["parent"])
        if n["level"] == 0:
            assert p == -1 and b[k]["spacing"] == np.float32(1024.0 / 128.0)
        else:
            assert int(b[p]["level"]) == n["level"] - 1
            assert (boxes[p][:3] <= boxes[k][:3]).all() and (boxes[k][3:] <= boxes[p][3:]).all()
            assert b[k]["spacing"] == b[p]["spacing"] / 2


def test_metadata_fields():
    meta, hier, octree, nodes = make_potree(seed=5, extra_attr=4)
    info = pcr.potree_info(meta)
    m = json.loads(meta)
    assert info.bytes_per_point == 12 + 4 + 6 and info.rgb_offset == 16
    assert info.num_points == m["points"] == sum(n["count"] for n in nodes)
    assert info.first_chunk_size == m["hierarchy"]["firstChunkSize"]
    m["encoding"] = "BROTLI"
    with pytest.raises(ValueError):
        pcr.potree_info(json.dumps(m))


def test_malformed_hierarchies_are_refused():
    meta, hier, octree, nodes = make_potree(seed=6)
    info = pcr.potree_info(meta)
    assert orc.potree_nodes(info, hier[:-22]) is None           # a chunk cut short
    info.first_chunk_size += 5                                   # not a multiple of 22
    assert orc.potree_nodes(info, hier) is None
    info = pcr.potree_info(meta)
    info.first_chunk_size -= 22                                  # children without records
    assert orc.potree_nodes(info, hier) is None


def test_encoding_round_trip_and_colour():
    meta, hier, octree, nodes = make_potree(seed=7)
    info = pcr.potree_info(meta)
    b, boxes = orc.potree_nodes(info, hier)
    lo, med, hi, rgba = orc.potree_encode(info, octree, b, boxes)
    rec = np.frombuffer(octree, np.uint8).reshape(-1, info.bytes_per_point)
    X = rec[:, 0:12].copy().view(np.int32).reshape(-1, 3).astype(np.float64)
    pos = X * np.array(info.scale) + np.array(info.offset) - np.array(info.min)
    for k in range(len(b)):
        f, n = int(b[k]["first_point"]), int(b[k]["num_points"])
        q = np.stack([((lo[f:f + n] >> (10 * a)) & 1023).astype(np.int64) << 20 |
                      ((med[f:f + n] >> (10 * a)) & 1023).astype(np.int64) << 10 |
                      ((hi[f:f + n] >> (10 * a)) & 1023).astype(np.int64) for a in range(3)], 1)
        size = boxes[k][3:] - boxes[k][:3]
        dec = boxes[k][:3] + q * (size / 2**30)
        inside = np.clip(pos[f:f + n], boxes[k][:3], boxes[k][3:])
        assert (np.abs(dec - inside) <= size * 2**-29).all()
    rgb = rec[:, info.rgb_offset:info.rgb_offset + 6].copy().view(np.uint16).reshape(-1, 3).astype(np.uint32)
    conv = np.where(rgb < 256, rgb, rgb // 256)
    assert np.array_equal(rgba, conv[:, 0] | (conv[:, 1] << 8) | (conv[:, 2] << 16))
