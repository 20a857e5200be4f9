"""CPU: LAS header parsing (C ABI and oracle) and the oracle's LAS encoder
(LasLoaderSparse.cpp:37-212, 495-526), pinned by the specification offsets,
the reference's example file (libs/laszip/example/5points_14.las, a LAS 1.4
point-format-6 file with 5 points), closed-form colour conversion and the
quantisation round trip."""
import ctypes as C
import os
import struct

import numpy as np
import pytest

import compute_rasterizer_amd as pcr
from compute_rasterizer_amd import capi
from oracle import orc
from compute_rasterizer_amd.datasets import make_las

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "5points_14.las")


def capi_parse(data):
    buf = np.frombuffer(data, np.uint8)
    info = capi.LasInfo()
    rc = capi.load().pcr_las_parse_header(capi.ptr(buf), buf.size, C.byref(info))
    return rc, info


def fields(info):
    return (info.num_points, info.offset_to_points, info.record_length, info.point_format, info.version_major,
            info.version_minor, info.rgb_offset, tuple(info.scale), tuple(info.offset), tuple(info.min),
            tuple(info.max))


def test_example_file_header():
    data = open(GOLDEN, "rb").read()
    rc, info = capi_parse(data)
    assert rc == 0
    # values read with struct at the LAS 1.4 specification offsets
    assert (info.version_major, info.version_minor) == (1, 4)
    assert info.num_points == struct.unpack_from("<Q", data, 247)[0] == 5
    assert info.offset_to_points == 921 and info.record_length == 33 and info.point_format == 6
    assert info.rgb_offset == 0          # format 6 has no colour; the reference reads bytes 0..5
    assert tuple(info.scale) == (0.01, 0.01, 0.01) and tuple(info.offset) == (600000.0, 4800000.0, 0.0)
    assert tuple(info.min) == (630498.56, 4834748.73, 61.33) and tuple(info.max) == (630499.95, 4834749.66, 63.68)
    assert fields(orc.las_parse_header(data)) == fields(info)


@pytest.mark.parametrize("fmt,version", [(2, (1, 2)), (3, (1, 3)), (7, (1, 4)), (8, (1, 4)), (1, (1, 2))])
def test_synthetic_headers(fmt, version):
    data, *_ = make_las(1001, point_format=fmt, version=version, seed=fmt)
    rc, info = capi_parse(data)
    assert rc == 0
    assert info.num_points == 1001 and info.point_format == fmt and (info.version_major, info.version_minor) == version
    assert info.rgb_offset == {2: 20, 3: 28, 7: 30, 8: 30}.get(fmt, 0)
    assert fields(orc.las_parse_header(data)) == fields(info)


def test_bad_headers_are_refused():
    data, *_ = make_las(100)
    assert capi_parse(b"LASX" + data[4:])[0] == capi.PCR_E_INVALID
    assert orc.las_parse_header(b"LASX" + data[4:]) is None
    assert capi_parse(data[:200])[0] == capi.PCR_E_INVALID            # shorter than a header
    assert capi_parse(data[:-5])[0] == capi.PCR_E_INVALID             # records cut short
    assert orc.las_parse_header(data[:-5]) is None
    compressed = bytearray(data)
    compressed[104] |= 128                                            # the LAZ bit is masked off
    assert capi_parse(bytes(compressed))[1].point_format == 2


def decode30(b, lo, med, hi):
    """float positions from the 30-bit planes (double arithmetic, no rounding to float)."""
    out = np.zeros((lo.size, 3))
    for k in range(len(b)):
        f, n = int(b[k]["first_point"]), int(b[k]["num_points"])
        for a in range(3):
            q = (((lo[f:f + n] >> (10 * a)) & 1023).astype(np.int64) << 20) | \
                (((med[f:f + n] >> (10 * a)) & 1023).astype(np.int64) << 10) | ((hi[f:f + n] >> (10 * a)) & 1023)
            size = float(b[k]["max"][a]) - float(b[k]["min"][a])
            out[f:f + n, a] = float(b[k]["min"][a]) + q * (size / 2**30)
    return out


def test_encode_round_trip_and_layout():
    n = 3 * 10240 + 77
    data, X, Y, Z, R, G, B = make_las(n, point_format=3, version=(1, 3), seed=5)
    info = orc.las_parse_header(data)
    b, lo, med, hi, rgba = orc.las_encode(data, info, sparse_offset=1001, file_index=2)
    assert len(b) == 4
    assert [int(x) for x in b["first_point"]] == [1001, 1001 + 10240, 1001 + 20480, 1001 + 30720]
    assert [int(x) for x in b["num_points"]] == [10240, 10240, 10240, 77]
    assert (b["level"] == 2).all() and (b["parent"] == -1).all()
    # positions relative to the header min, in double
    pos = np.stack([X * info.scale[0] + info.offset[0] - info.min[0], Y * info.scale[1] + info.offset[1] - info.min[1],
                    Z * info.scale[2] + info.offset[2] - info.min[2]], 1)
    b0 = b.copy()
    b0["first_point"] -= 1001
    dec = decode30(b0, lo, med, hi)
    # quantisation error below one 30-bit step plus the float rounding of the header
    for k in range(len(b0)):
        f, m = int(b0[k]["first_point"]), int(b0[k]["num_points"])
        seg = pos[f:f + m]
        span = seg.max(0) - seg.min(0)
        tol = span * 2**-30 * 2 + np.abs(seg).max(0) * 2**-23 + 1e-9
        assert (np.abs(dec[f:f + m] - seg) <= tol).all()
        # the batch AABB is the float of the double min/max
        assert np.array_equal(b0[k]["min"], seg.min(0).astype(np.float32))
        assert np.array_equal(b0[k]["max"], seg.max(0).astype(np.float32))


def test_quantisation_corners_and_colour():
    data, X, Y, Z, R, G, B = make_las(10240, point_format=2, seed=9)
    info = orc.las_parse_header(data)
    b, lo, med, hi, rgba = orc.las_encode(data, info)
    pos = np.stack([X.astype(np.float64), Y, Z], 1)
    imin, imax = pos.argmin(0), pos.argmax(0)
    for a in range(3):
        # the point at the batch min quantises to 0, the point at the max to 2^30 - 1
        q = lambda i: (int(lo[i] >> (10 * a)) & 1023, int(med[i] >> (10 * a)) & 1023, int(hi[i] >> (10 * a)) & 1023)
        assert q(imin[a]) == (0, 0, 0)
        assert q(imax[a]) == (1023, 1023, 1023)
    conv = lambda v: np.where(v < 256, v, v // 256).astype(np.uint32)
    assert np.array_equal(rgba, conv(R) | (conv(G) << 8) | (conv(B) << 16))
    # closed-form colour cases
    for v, want in [(0, 0), (255, 255), (256, 1), (300, 1), (65535, 255), (511, 1), (512, 2)]:
        assert int(conv(np.array([v]))[0]) == want


def test_flat_axis_quantises_to_zero():
    data, *_ = make_las(5000, flat_z=True, seed=3)
    info = orc.las_parse_header(data)
    b, lo, med, hi, rgba = orc.las_encode(data, info)
    assert ((lo >> 20) & 1023 == 0).all() and ((med >> 20) & 1023 == 0).all() and ((hi >> 20) & 1023 == 0).all()
    assert (b["min"][:, 2] == b["max"][:, 2]).all()


def test_example_file_encodes():
    data = open(GOLDEN, "rb").read()
    info = orc.las_parse_header(data)
    b, lo, med, hi, rgba = orc.las_encode(data, info)
    assert len(b) == 1 and int(b[0]["num_points"]) == 5
    # the batch AABB is the float of the records' double box relative to the header min
    xyz = np.array([[struct.unpack_from("<i", data, 921 + 33 * i + 4 * a)[0] * info.scale[a] + info.offset[a]
                     - info.min[a] for a in range(3)] for i in range(5)])
    assert np.array_equal(b[0]["min"], xyz.min(0).astype(np.float32))
    assert np.array_equal(b[0]["max"], xyz.max(0).astype(np.float32))
    # format 6 has no colour: the reference's offset 0 reads the low bytes of X and Y
    X = [struct.unpack_from("<HHH", data, 921 + 33 * i) for i in range(5)]
    conv = lambda v: v if v < 256 else v // 256
    assert [int(c) for c in rgba] == [conv(r) | (conv(g) << 8) | (conv(bb) << 16) for r, g, bb in X]
