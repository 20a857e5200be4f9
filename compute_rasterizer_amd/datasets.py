"""Synthetic LAS and Potree 2.0 datasets (used by the ingest tests and the bench).

make_las writes the public LAS 1.2-1.4 layout: the header fields at their
specification offsets, no VLRs, and point records of the format's length
with X, Y, Z as int32 at bytes 0..11 and R, G, B as uint16 at the format's
colour offset; the remaining record bytes are seeded noise.

make_potree writes the published Potree 2.0 layout: metadata.json,
hierarchy.bin (22-byte node records: type, child mask, point count, byte
offset, byte size; each chunk lists its subtree breadth first, and nodes at
the chunk's depth limit are proxies of type 2 whose offset and size point at
their own chunk) and octree.bin (each node's point records contiguous). The
root box is a cube of 1024 m, so every child box is exact in float."""
import json
import struct

import numpy as np

RECORD_LENGTH = {0: 20, 1: 28, 2: 26, 3: 34, 6: 30, 7: 36, 8: 38}
RGB_OFFSET = {2: 20, 3: 28, 5: 28, 7: 30, 8: 30, 10: 30}


def make_las(n, point_format=2, version=(1, 2), seed=0, scale=(0.001, 0.001, 0.001),
             offset=(500000.0, 4000000.0, 100.0), extent=(300.0, 200.0, 40.0), rgb16=True, flat_z=False,
             extra_bytes=0):
    """(bytes, X, Y, Z, R, G, B) of an n-point LAS file."""
    rng = np.random.default_rng(seed)
    rl = RECORD_LENGTH[point_format] + extra_bytes
    hdr_size = 375 if version >= (1, 4) else 227
    X = rng.integers(0, int(extent[0] / scale[0]), n).astype(np.int32)
    Y = rng.integers(0, int(extent[1] / scale[1]), n).astype(np.int32)
    Z = np.full(n, 1234, np.int32) if flat_z else rng.integers(0, int(extent[2] / scale[2]), n).astype(np.int32)
    if rgb16:
        R, G, B = (rng.integers(0, 65536, n).astype(np.uint16) for _ in range(3))
    else:
        R, G, B = (rng.integers(0, 256, n).astype(np.uint16) for _ in range(3))
    rec = np.frombuffer(rng.bytes(n * rl), np.uint8).reshape(n, rl).copy() if n else np.zeros((0, rl), np.uint8)
    rec[:, 0:4] = X.view(np.uint8).reshape(n, 4)
    rec[:, 4:8] = Y.view(np.uint8).reshape(n, 4)
    rec[:, 8:12] = Z.view(np.uint8).reshape(n, 4)
    if point_format in RGB_OFFSET:
        o = RGB_OFFSET[point_format]
        for k, c in enumerate((R, G, B)):
            rec[:, o + 2 * k:o + 2 * k + 2] = c.view(np.uint8).reshape(n, 2)
    xs = [X * scale[0] + offset[0], Y * scale[1] + offset[1], Z * scale[2] + offset[2]]
    mins = [float(v.min()) if n else 0.0 for v in xs]
    maxs = [float(v.max()) if n else 0.0 for v in xs]
    h = bytearray(hdr_size)
    h[0:4] = b"LASF"
    h[24], h[25] = version
    struct.pack_into("<H", h, 94, hdr_size)
    struct.pack_into("<I", h, 96, hdr_size)          # offset to point data (no VLRs)
    struct.pack_into("<I", h, 100, 0)
    h[104] = point_format
    struct.pack_into("<H", h, 105, rl)
    struct.pack_into("<I", h, 107, n if (version < (1, 4) or n < 2**32) else 0)
    struct.pack_into("<3d", h, 131, *scale)
    struct.pack_into("<3d", h, 155, *offset)
    struct.pack_into("<6d", h, 179, maxs[0], mins[0], maxs[1], mins[1], maxs[2], mins[2])
    if version >= (1, 4):
        struct.pack_into("<Q", h, 247, n)
    return bytes(h) + rec.tobytes(), X, Y, Z, R, G, B


ROOT = 1024.0


def make_potree(seed=0, max_depth=4, chunk_depth=2, bbmin=(10000.0, 20000.0, 300.0), scale=0.001,
                offset=(10000.0, 20000.0, 0.0), extra_attr=2, shuffle_octree=True, points_per_node=(20, 300)):
    """(metadata_text, hierarchy_bytes, octree_bytes, nodes) with nodes a list of dicts
    {name, box (relative min+max), level, count, first (point index)}."""
    rng = np.random.default_rng(seed)
    bpp = 12 + extra_attr + 6
    rgb_off = 12 + extra_attr
    nodes = []

    def make(name, box, level):
        n = {"name": name, "box": box, "level": level, "children": [None] * 8,
             "count": int(rng.integers(points_per_node[0], points_per_node[1]))}
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
