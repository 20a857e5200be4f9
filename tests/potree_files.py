"""Synthetic Potree 2.0 datasets for the ingest tests (test helper).

Writes the published Potree 2.0 layout: metadata.json, hierarchy.bin (22-byte
node records: type, child mask, point count, byte offset, byte size; each
chunk lists its subtree breadth first, and nodes at the chunk's depth limit
are proxies of type 2 whose offset and size point at their own chunk) and
octree.bin (each node's point records contiguous). The root box is a cube
of 1024 m, so every child box is exact in float."""
import json
import struct

import numpy as np

ROOT = 1024.0


def make_potree(seed=0, max_depth=4, chunk_depth=2, bbmin=(10000.0, 20000.0, 300.0), scale=0.001,
                offset=(10000.0, 20000.0, 0.0), extra_attr=2, shuffle_octree=True):
    """(metadata_text, hierarchy_bytes, octree_bytes, nodes) with nodes a list of dicts
    {name, box (relative min+max), level, count, first (point index)}."""
    rng = np.random.default_rng(seed)
    bpp = 12 + extra_attr + 6
    rgb_off = 12 + extra_attr
    nodes = []

    def make(name, box, level):
        n = {"name": name, "box": box, "level": level, "children": [None] * 8,
             "count": int(rng.integers(20, 300))}
        nodes.append(n)
        if level < max_depth:
            for k in range(8):
                if rng.random() < (0.9 if level < 2 else 0.35):
                    mn, mx = list(box[:3]), list(box[3:])
                    for bit in range(3):
                        a = 2 - bit
                        half = (mx[a] - mn[a]) / 2
                        if k & (1 << bit):
                            mn[a] += half
                        else:
                            mx[a] -= half
                    n["children"][k] = make(name + str(k), tuple(mn + mx), level + 1)
        return n

    root = make("r", (0.0, 0.0, 0.0, ROOT, ROOT, ROOT), 0)

    # point records, nodes laid out in octree.bin in a shuffled order
    order = list(range(len(nodes)))
    if shuffle_octree:
        rng.shuffle(order)
    recs, first = [], 0
    for i in order:
        n = nodes[i]
        c = n["count"]
        lo, hi = np.array(n["box"][:3]), np.array(n["box"][3:])
        rel = lo + rng.random((c, 3)) * (hi - lo)
        world = rel + np.array(bbmin)
        X = np.round((world - np.array(offset)) / scale).astype(np.int32)
        rec = np.frombuffer(rng.bytes(c * bpp), np.uint8).reshape(c, bpp).copy()
        rec[:, 0:12] = X.view(np.uint8).reshape(c, 12)
        rgb = rng.integers(0, 65536, (c, 3)).astype(np.uint16)
        rec[:, rgb_off:rgb_off + 6] = rgb.view(np.uint8).reshape(c, 6)
        n["first"] = first
        first += c
        recs.append(rec)
    octree = np.concatenate(recs).tobytes()
    npoints = first

    # hierarchy: breadth-first chunks of chunk_depth levels, proxies below
    chunks = []   # list of bytearrays, chunk 0 first

    def chunk_records(top):
        """records of top's chunk; returns the bytearray and the proxies to fill later."""
        out, proxies, queue = bytearray(), [], [(top, 0)]
        while queue:
            n, d = queue.pop(0)
            mask = sum(1 << k for k in range(8) if n["children"][k] is not None)
            if d == chunk_depth and n is not top and mask:
                proxies.append((len(out), n))
                out += struct.pack("<BBIqq", 2, mask, n["count"], 0, 0)
                continue
            typ = 1 if mask == 0 else 0
            out += struct.pack("<BBIqq", typ, mask, n["count"], n["first"] * bpp, n["count"] * bpp)
            queue += [(ch, d + 1) for ch in n["children"] if ch is not None]
        return out, proxies

    hier = bytearray()
    pending = [(None, root)]
    while pending:
        patch, top = pending.pop(0)
        recs_, proxies = chunk_records(top)
        off = len(hier)
        if patch is not None:
            struct.pack_into("<qq", hier, patch + 6, off, len(recs_))
        hier += recs_
        pending += [(off + pos, n) for pos, n in proxies]
    first_chunk = len(chunk_records(root)[0])

    bbmax = [bbmin[a] + ROOT for a in range(3)]
    meta = {"version": "2.0", "name": "synthetic", "description": "", "points": npoints, "projection": "",
            "hierarchy": {"firstChunkSize": first_chunk, "stepSize": chunk_depth, "depth": max_depth},
            "offset": list(offset), "scale": [scale] * 3, "spacing": ROOT / 128.0,
            "boundingBox": {"min": list(bbmin), "max": bbmax}, "encoding": "DEFAULT",
            "attributes": [{"name": "position", "description": "", "size": 12, "numElements": 3, "elementSize": 4,
                            "type": "int32"},
                           {"name": "intensity", "description": "", "size": extra_attr, "numElements": 1,
                            "elementSize": extra_attr, "type": "uint16"},
                           {"name": "rgb", "description": "", "size": 6, "numElements": 3, "elementSize": 2,
                            "type": "uint16"}]}
    return json.dumps(meta), bytes(hier), octree, nodes
