"""Writes the LAZ fixtures under tests/golden/laz/ and their expected records.

Run in the build container (the reference's sources must be under
/root/reference): `python tests/golden/make_laz.py`. It

1. builds the reference's vendored LASzip into oracle/_ref/laz_ref_tool
   (`make -C oracle ref`),
2. writes seeded LAS files of point formats 0-3 and 6-8, with structured
   fields (scan lines, pulses of 1-15 returns, GPS time sequences with
   jumps and interleaving, grey and 8-bit colours, scanner channel
   switches, classes above 31, extra bytes) and a noise case,
3. compresses each with the reference's LASzip (fixed, variable and
   default chunk sizes; LAS 1.4 items as version 3 and version 4),
4. checks that the reference's LASzip decodes each file back to the LAS
   records, and
5. stores the .laz files and, in laz_records.npz, the LAS header and records
   each must decode to (LAZ is lossless).
"""
import os
import struct
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
TOOL = os.path.join(ROOT, "oracle", "_ref", "laz_ref_tool")
OUT_DIR = os.path.join(HERE, "laz")
STD_LEN = {0: 20, 1: 28, 2: 26, 3: 34, 6: 30, 7: 36, 8: 38}


def structured_records(n, fmt, extra, seed):
    """n records of point format fmt (+ extra bytes) with realistic structure."""
    rng = np.random.default_rng(seed)
    rl = STD_LEN[fmt] + extra
    rec = np.zeros((n, rl), np.uint8)
    v14 = fmt >= 6
    maxret = 15 if v14 else 7
    # pulses of 1..maxret returns
    r = np.zeros(n, np.int64)
    nr = np.zeros(n, np.int64)
    pulse = np.zeros(n, np.int64)
    i = p = 0
    while i < n:
        k = int(rng.choice([1, 1, 1, 2, 2, 3, 4, 5, maxret]))
        k = min(k, n - i)
        nr[i:i + k] = k
        r[i:i + k] = np.arange(1, k + 1)
        pulse[i:i + k] = p
        i += k
        p += 1
    odd = rng.random(n) < 0.01           # damaged return fields
    r[odd] = rng.integers(0, maxret + 1, odd.sum())
    nr[odd] = rng.integers(0, maxret + 1, odd.sum())
    # positions: scan lines of ~200 pulses, returns below the first
    line = pulse // 200
    x = (pulse % 200) * 137 + rng.integers(-20, 21, n) + line * 7
    y = line * 450 + rng.integers(-30, 31, n)
    z = 50_000 + np.cumsum(rng.integers(-40, 41, n)) - (r - 1) * rng.integers(200, 3000, n)
    jump = rng.random(n) < 0.002
    x = x + np.cumsum(jump * rng.integers(-10**7, 10**7, n))
    for k, v in enumerate((x, y, z)):
        rec[:, 4 * k:4 * k + 4] = v.astype(np.int32).view(np.uint8).reshape(n, 4)
    inten = np.clip(np.cumsum(rng.integers(-300, 301, n)) % 65536, 0, 65535).astype(np.uint16)
    inten[rng.random(n) < 0.05] = 0
    rec[:, 12:14] = inten.view(np.uint8).reshape(n, 2)
    scan_dir = (line % 2).astype(np.uint8)
    edge = ((pulse % 200) == 199).astype(np.uint8)
    cls = rng.choice([1, 2, 2, 2, 3, 5, 6, 9], n).astype(np.uint8)
    if v14:
        big = rng.random(n) < 0.03
        cls[big] = rng.integers(32, 256, big.sum())
    user = np.where(rng.random(n) < 0.9, 7, rng.integers(0, 256, n)).astype(np.uint8)
    psid = (100 + line // 5).astype(np.uint16)
    angle = ((pulse % 200) - 100) // 7
    if not v14:
        rec[:, 14] = (np.minimum(r, 7) | (np.minimum(nr, 7) << 3) | (scan_dir << 6) | (edge << 7)).astype(np.uint8)
        rec[:, 15] = cls | (((rng.random(n) < 0.02) * rng.integers(1, 8, n)) << 5).astype(np.uint8)
        rec[:, 16] = angle.astype(np.int8).view(np.uint8)
        rec[:, 17] = user
        rec[:, 18:20] = psid.view(np.uint8).reshape(n, 2)
        o = 20
    else:
        flags = ((rng.random(n) < 0.03) * rng.integers(1, 16, n)).astype(np.uint8)
        ch = np.zeros(n, np.int64)
        c, i = 0, 0
        while i < n:   # bursts on one scanner channel
            k = int(rng.integers(1, 60))
            ch[i:i + k] = c
            i += k
            c = int(rng.integers(0, 4))
        rec[:, 14] = (r | (nr << 4)).astype(np.uint8)
        rec[:, 15] = (flags | (ch << 4) | (scan_dir << 6) | (edge << 7)).astype(np.uint8)
        rec[:, 16] = cls
        rec[:, 17] = user
        sa = (angle * 166).astype(np.int16)
        wild = rng.random(n) < 0.01
        sa[wild] = rng.integers(-30000, 30000, wild.sum())
        rec[:, 18:20] = sa.view(np.uint8).reshape(n, 2)
        rec[:, 20:22] = psid.view(np.uint8).reshape(n, 2)
        o = 22
    if fmt in (1, 3, 6, 7, 8):   # GPS time: per pulse, two interleaved flight lines, jumps
        base = np.where(pulse % 97 < 80, 1.6e8, 3.2e8)
        t = base + pulse * 1.5e-5 + (line // 50) * 12.5
        t[rng.random(n) < 0.001] += 1e6
        same = rng.random(n) < 0.02
        t[same] = np.roll(t, 1)[same]
        rec[:, o:o + 8] = t.astype(np.float64).view(np.uint8).reshape(n, 8)
        o += 8
    if fmt in (2, 3, 7, 8):
        base = np.cumsum(rng.integers(-3, 4, (n, 3)), axis=0) % 256
        rgb = base.astype(np.uint16)
        grey = rng.random(n) < 0.2
        rgb[grey, 1] = rgb[grey, 0]
        rgb[grey, 2] = rgb[grey, 0]
        wide = (np.arange(n) // 500) % 2 == 1        # 16-bit colours in every other block
        rgb[wide] = rgb[wide] * 257 + rng.integers(0, 3, (wide.sum(), 3)).astype(np.uint16)
        rec[:, o:o + 6] = rgb.view(np.uint8).reshape(n, 6)
        o += 6
        if fmt == 8:
            nir = (np.cumsum(rng.integers(-50, 51, n)) % 65536).astype(np.uint16)
            rec[:, o:o + 2] = nir.view(np.uint8).reshape(n, 2)
            o += 2
    if extra:
        ex = np.zeros((n, extra), np.uint8)
        ex[:, 0] = 42
        if extra > 1:
            ex[:, 1] = np.arange(n) % 256
        if extra > 2:
            ex[:, 2:] = rng.integers(0, 256, (n, extra - 2))
        rec[:, o:o + extra] = ex
        o += extra
    assert o == rl, (o, rl)
    return rec


def las_bytes(rec, fmt, version):
    n, rl = rec.shape
    hdr = 375 if version >= (1, 4) else 227
    scale, offset = (0.001, 0.001, 0.001), (500000.0, 4000000.0, 0.0)
    xyz = rec[:, :12].copy().view(np.int32).reshape(n, 3).astype(np.float64) * np.array(scale) + np.array(offset)
    h = bytearray(hdr)
    h[0:4] = b"LASF"
    h[24], h[25] = version
    struct.pack_into("<H", h, 94, hdr)
    struct.pack_into("<I", h, 96, hdr)
    h[104] = fmt
    struct.pack_into("<H", h, 105, rl)
    struct.pack_into("<I", h, 107, n)
    struct.pack_into("<3d", h, 131, *scale)
    struct.pack_into("<3d", h, 155, *offset)
    mn, mx = xyz.min(0), xyz.max(0)
    struct.pack_into("<6d", h, 179, mx[0], mn[0], mx[1], mn[1], mx[2], mn[2])
    if version >= (1, 4):
        struct.pack_into("<Q", h, 247, n)
    return bytes(h) + rec.tobytes()


# name: (format, version, points, chunk argument, v4, extra bytes, seed, noise)
CASES = {
    "f0_chunk1000": (0, (1, 2), 3100, "1000", False, 0, 10, False),
    "f1_gps": (1, (1, 2), 4000, None, False, 0, 11, False),
    "f2_rgb_chunk1500": (2, (1, 2), 4100, "1500", False, 0, 12, False),
    "f3_extra3_variable": (3, (1, 3), 5000, "variable:777", False, 3, 13, False),
    "f3_noise_chunk700": (3, (1, 2), 2000, "700", False, 0, 14, True),
    "f6_channels": (6, (1, 4), 4000, "1200", False, 0, 15, False),
    "f7_v3": (7, (1, 4), 5000, "1000", False, 0, 16, False),
    "f7_v4": (7, (1, 4), 5000, "1000", True, 0, 16, False),
    "f8_extra2_v3": (8, (1, 4), 4500, "2000", False, 2, 17, False),
    "f8_extra2_v4": (8, (1, 4), 4500, "2000", True, 2, 17, False),
}


def main():
    subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    os.makedirs(OUT_DIR, exist_ok=True)
    store = {}
    with tempfile.TemporaryDirectory() as tmp:
        for name, (fmt, ver, n, chunk, v4, extra, seed, noise) in CASES.items():
            if noise:
                rng = np.random.default_rng(seed)
                rec = np.frombuffer(rng.bytes(n * (STD_LEN[fmt] + extra)), np.uint8).reshape(n, -1).copy()
            else:
                rec = structured_records(n, fmt, extra, seed)
            las = las_bytes(rec, fmt, ver)
            src, laz, back = (os.path.join(tmp, name + s) for s in (".las", ".laz", ".bin"))
            open(src, "wb").write(las)
            cmd = [TOOL, "compress", src, laz] + ([chunk] if chunk or v4 else []) + (["v4"] if v4 else [])
            if v4 and not chunk:
                cmd[4] = "50000"
            subprocess.run(cmd, check=True)
            subprocess.run([TOOL, "decompress", laz, back], check=True)
            got = np.fromfile(back, np.uint8).reshape(rec.shape)
            assert np.array_equal(got, rec), f"{name}: the reference's LASzip does not round-trip"
            data = open(laz, "rb").read()
            open(os.path.join(OUT_DIR, name + ".laz"), "wb").write(data)
            key = name[:-3] + "_v3" if name.endswith("_v4") else name
            store[key + "_records"] = rec
            store[key + "_las_header"] = np.frombuffer(las[:(375 if ver >= (1, 4) else 227)], np.uint8)
            print(f"{name}: {n} points, LAS {len(las)} B -> LAZ {len(data)} B", flush=True)
    np.savez_compressed(os.path.join(HERE, "laz_records.npz"), **store)


if __name__ == "__main__":
    sys.exit(main())
