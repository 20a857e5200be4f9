"""Synthetic uncompressed LAS files for the ingest tests (test helper).

Writes the public LAS 1.2-1.4 layout: the header fields at their
specification offsets, no VLRs, and point records of the format's length
with X, Y, Z as int32 at bytes 0..11 and R, G, B as uint16 at the format's
colour offset. The remaining record bytes are seeded noise."""
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
