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


# ---------------------------------------------------------------- LAZ tiling
# A large LAZ file for throughput runs, made from a small one by repeating its
# chunks (each chunk restarts every model, so a chunk decodes the same
# anywhere) behind a new chunk table. The table is arithmetic coded: an
# integer compressor (32 bits, 2 contexts) codes each chunk's point count
# (context 0) and byte size (context 1) as a difference from the previous
# chunk's. The encoder below is the inverse of the decoder in
# compute_rasterizer_amd/csrc/pcr_laz.cpp (A. Said's FastAC coder as LASzip
# uses it).

class _Model:
    def __init__(self, n):
        self.n, self.cnt, self.total, self.cycle = n, [1] * n, 0, n
        self.dist = [0] * n
        self._update()
        self.cycle = self.until = (n + 6) >> 1

    def _update(self):
        self.total += self.cycle
        if self.total > (1 << 15):
            self.cnt = [(c + 1) >> 1 for c in self.cnt]
            self.total = sum(self.cnt)
        scale, acc = 0x80000000 // self.total, 0
        for k in range(self.n):
            self.dist[k] = ((scale * acc) & 0xFFFFFFFF) >> 16
            acc += self.cnt[k]
        self.cycle = min((5 * self.cycle) >> 2, (self.n + 6) << 3)
        self.until = self.cycle


class _BitModel:
    def __init__(self):
        self.p0, self.n0, self.n, self.cycle, self.until = 1 << 12, 1, 2, 4, 4

    def _update(self):
        self.n += self.cycle
        if self.n > (1 << 13):
            self.n, self.n0 = (self.n + 1) >> 1, (self.n0 + 1) >> 1
            if self.n0 == self.n:
                self.n += 1
        self.p0 = (self.n0 * (0x80000000 // self.n)) >> 18
        self.cycle = min((5 * self.cycle) >> 2, 64)
        self.until = self.cycle


class _Encoder:
    MIN = 1 << 24

    def __init__(self):
        self.base, self.length, self.out = 0, 0xFFFFFFFF, bytearray()

    def _carry(self):
        i = len(self.out) - 1
        while self.out[i] == 0xFF:
            self.out[i] = 0
            i -= 1
        self.out[i] += 1

    def _step(self, x, length):
        init = self.base
        self.base = (self.base + x) & 0xFFFFFFFF
        self.length = length
        if init > self.base:
            self._carry()
        while self.length < self.MIN:
            self.out.append(self.base >> 24)
            self.base = (self.base << 8) & 0xFFFFFFFF
            self.length = (self.length << 8) & 0xFFFFFFFF

    def symbol(self, m, s):
        if s == m.n - 1:
            x = m.dist[s] * (self.length >> 15)
            self._step(x, self.length - x)
        else:
            ln = self.length >> 15
            x = m.dist[s] * ln
            self._step(x, m.dist[s + 1] * ln - x)
        m.cnt[s] += 1
        m.until -= 1
        if m.until == 0:
            m._update()

    def bit(self, m, b):
        x = m.p0 * (self.length >> 13)
        if b == 0:
            m.n0 += 1
            self._step(0, x)
        else:
            self._step(x, self.length - x)
        m.until -= 1
        if m.until == 0:
            m._update()

    def bits(self, n, v):
        if n > 19:
            self.bits(16, v & 0xFFFF)
            v, n = v >> 16, n - 16
        ln = self.length >> n
        self._step(v * ln, ln)

    def done(self):
        init = self.base
        another = self.length > 2 * self.MIN
        if another:
            self.base, self.length = (self.base + self.MIN) & 0xFFFFFFFF, self.MIN >> 1
        else:
            self.base, self.length = (self.base + (self.MIN >> 1)) & 0xFFFFFFFF, self.MIN >> 9
        if init > self.base:
            self._carry()
        while self.length < self.MIN:
            self.out.append(self.base >> 24)
            self.base = (self.base << 8) & 0xFFFFFFFF
            self.length = (self.length << 8) & 0xFFFFFFFF
        self.out += b"\0\0\0" if another else b"\0\0"
        return bytes(self.out)


def _encode_chunk_table(counts, sizes):
    """The arithmetic-coded body of a LAZ chunk table (variable chunk sizes)."""
    enc = _Encoder()
    k_models = [_Model(33), _Model(33)]
    c0 = _BitModel()
    corr = [None] + [_Model(1 << min(i, 8)) for i in range(1, 33)]

    def put(pred, real, ctx):
        c = (real - pred) & 0xFFFFFFFF
        c = c - (1 << 32) if c >= 1 << 31 else c
        c1 = -c if c <= 0 else c - 1
        k = c1.bit_length()
        enc.symbol(k_models[ctx], k)
        if k == 0:
            enc.bit(c0, c)
        elif k < 32:
            c = c + (1 << k) - 1 if c < 0 else c - 1
            if k <= 8:
                enc.symbol(corr[k], c)
            else:
                low = k - 8
                enc.symbol(corr[k], c >> low)
                enc.bits(low, c & ((1 << low) - 1))

    pc = pb = 0
    for n, b in zip(counts, sizes):
        put(pc, int(n), 0)
        put(pb, int(b), 1)
        pc, pb = int(n), int(b)
    return enc.done()


def laz_tile(data: bytes, copies: int) -> bytes:
    """A LAZ file whose records are `copies` repetitions of data's records: its
    chunks repeated behind a new (variable-size) chunk table."""
    from . import laz_chunk_table
    starts, firsts = laz_chunk_table(data)
    assert len(starts) > 1, "the file has no chunk table"
    n = int(firsts[-1])
    p0 = struct.unpack_from("<I", data, 96)[0]
    chunks = data[int(starts[0]):int(starts[-1])]
    counts = np.diff(firsts).astype(np.int64)
    sizes = np.diff(starts).astype(np.int64)
    head = bytearray(data[:p0])
    total = n * copies
    if head[24] == 1 and head[25] < 4:
        assert total < 2**32
        struct.pack_into("<I", head, 107, total)
    else:
        struct.pack_into("<I", head, 107, total if total < 2**32 else 0)
        struct.pack_into("<Q", head, 247, total)
    # the LASzip VLR's chunk size becomes "variable"
    hs, nvlr = struct.unpack_from("<H", head, 94)[0], struct.unpack_from("<I", head, 100)[0]
    p = hs
    for _ in range(nvlr):
        rec, ln = struct.unpack_from("<HH", head, p + 18)
        if rec == 22204 and head[p + 2:p + 16] == b"laszip encoded":
            struct.pack_into("<I", head, p + 54 + 12, 0xFFFFFFFF)
        p += 54 + ln
    body = chunks * copies
    table_at = p0 + 8 + len(body)
    table = struct.pack("<II", 0, len(counts) * copies) + _encode_chunk_table(np.tile(counts, copies),
                                                                             np.tile(sizes, copies))
    return bytes(head) + struct.pack("<q", table_at) + body + table


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
