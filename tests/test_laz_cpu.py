"""CPU: the library's LAZ decoder against the fixtures the reference's vendored
LASzip wrote (tests/golden/make_laz.py).

LAZ is lossless, so each fixture must decode to the LAS records it was
compressed from, byte for byte: point formats 0-3 (pointwise chunked,
version 2 items) and 6-8 (layered chunked, version 3 and 4 items), fixed,
variable and single chunks, extra bytes, GPS time sequences, scanner channel
switches and a noise case. Where the build container has the reference's
sources, the reference's own LASzip decoder (oracle/_ref) must agree too.
Damaged files must fail cleanly, never crash.
"""
import ctypes as C
import os
import struct
import subprocess

import numpy as np
import pytest

import compute_rasterizer_amd as pcr
from compute_rasterizer_amd import capi

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")
LAZ_DIR = os.path.join(GOLDEN, "laz")
REF_TOOL = os.path.join(os.path.dirname(HERE), "oracle", "_ref", "laz_ref_tool")
CASES = sorted(f[:-4] for f in os.listdir(LAZ_DIR) if f.endswith(".laz"))
# item lists of the cases (LASzip item types: 6 POINT10, 7 GPSTIME11, 8 RGB12,
# 0 BYTE, 10 POINT14, 11 RGB14, 12 RGBNIR14, 14 BYTE14)
ITEMS = {"f0_chunk1000": [6], "f1_gps": [6, 7], "f2_rgb_chunk1500": [6, 8], "f3_extra3_variable": [6, 7, 8, 0],
         "f3_noise_chunk700": [6, 7, 8], "f6_channels": [10], "f7_v3": [10, 11], "f7_v4": [10, 11],
         "f8_extra2_v3": [10, 12, 14], "f8_extra2_v4": [10, 12, 14]}


@pytest.fixture(scope="module")
def records():
    return np.load(os.path.join(GOLDEN, "laz_records.npz"))


def laz(name):
    with open(os.path.join(LAZ_DIR, name + ".laz"), "rb") as f:
        return f.read()


def expected(records, name):
    key = name[:-3] + "_v3" if name.endswith("_v4") else name
    return records[key + "_records"], records[key + "_las_header"].tobytes()


def test_every_case_has_a_fixture():
    assert CASES == sorted(ITEMS)


@pytest.mark.parametrize("name", CASES)
def test_fixture_decodes_to_its_las_records(records, name):
    rec, _ = expected(records, name)
    got = pcr.laz_decode(laz(name))
    assert got.shape == rec.shape
    assert np.array_equal(got, rec), int((got != rec).any(1).sum())


@pytest.mark.parametrize("name", CASES)
def test_info(records, name):
    rec, hdr = expected(records, name)
    info = pcr.laz_info(laz(name))
    assert info.num_points == rec.shape[0] and info.record_length == rec.shape[1]
    assert info.point_format == hdr[104]
    assert [info.item_type[i] for i in range(info.num_items)] == ITEMS[name]
    fmt = hdr[104]
    assert info.compressor == (3 if fmt >= 6 else 2)
    want_version = 2 if fmt < 6 else (4 if name.endswith("_v4") else 3)
    assert all(info.item_version[i] == want_version for i in range(info.num_items))
    assert info.has_rgb == (fmt in (2, 3, 7, 8))
    assert info.rgb_offset == {2: 20, 3: 28, 7: 30, 8: 30}.get(fmt, 0)
    n = rec.shape[0]
    if name == "f3_extra3_variable":
        assert info.chunk_size == 0xFFFFFFFF and info.num_chunks == (n + 776) // 777
    elif name == "f1_gps":
        assert info.chunk_size == 50000 and info.num_chunks == 1
    else:
        assert info.num_chunks == (n + info.chunk_size - 1) // info.chunk_size


@pytest.mark.parametrize("name", ["f0_chunk1000", "f3_extra3_variable", "f7_v3", "f8_extra2_v4"])
def test_ranges_and_threads(records, name):
    rec, _ = expected(records, name)
    data, n = laz(name), rec.shape[0]
    for first, count, threads in ((0, 1, 1), (1, 1, 1), (17, 1234, 3), (999, 2, 2), (776, 778, 4), (n - 5, 5, 0),
                                  (n // 2, n - n // 2, 8), (0, n, 64), (n, 0, 1)):
        got = pcr.laz_decode(data, first, count, threads=threads)
        assert np.array_equal(got, rec[first:first + count]), (first, count, threads)


@pytest.mark.parametrize("name", CASES)
def test_xyz_rgb_fields(records, name):
    """Decoding only the position and colour layers gives the same X, Y, Z and R, G, B."""
    rec, _ = expected(records, name)
    info = pcr.laz_info(laz(name))
    got = pcr.laz_decode(laz(name), fields=capi.LAZ_XYZ_RGB)
    assert np.array_equal(got[:, :12], rec[:, :12])
    if info.has_rgb:
        o = info.rgb_offset
        assert np.array_equal(got[:, o:o + 6], rec[:, o:o + 6])


@pytest.mark.parametrize("name", CASES)
def test_las_header_marks_compressed(records, name):
    _, hdr = expected(records, name)
    data = laz(name)
    info = capi.LasInfo()
    buf = np.frombuffer(data, np.uint8)
    assert capi.load().pcr_las_parse_header(buf.ctypes.data, buf.size, C.byref(info)) == 0
    assert info.compressed == 1 and info.point_format == hdr[104]
    plain = capi.LasInfo()
    las = np.frombuffer(hdr, np.uint8)
    # the bare LAS header of the same file is not compressed (records absent, so only count 0 parses)
    h = bytearray(hdr)
    struct.pack_into("<I", h, 107, 0)
    if len(h) >= 255:
        struct.pack_into("<Q", h, 247, 0)
    las = np.frombuffer(bytes(h), np.uint8)
    assert capi.load().pcr_las_parse_header(las.ctypes.data, las.size, C.byref(plain)) == 0
    assert plain.compressed == 0


def _patch(data, off, fmt, *vals):
    b = bytearray(data)
    struct.pack_into(fmt, b, off, *vals)
    return bytes(b)


@pytest.mark.parametrize("name", ["f0_chunk1000", "f7_v3"])
def test_missing_chunk_table_decodes_in_order(records, name):
    """A writer that stopped before its chunk table leaves the pointer at the
    first chunk; the chunks are then decoded one after the other."""
    rec, _ = expected(records, name)
    data = laz(name)
    p0 = struct.unpack_from("<I", data, 96)[0]
    table = struct.unpack_from("<q", data, p0)[0]
    cut = _patch(data[:table], p0, "<q", p0)
    info = pcr.laz_info(cut)
    assert info.num_chunks == 0
    assert np.array_equal(pcr.laz_decode(cut), rec)
    assert np.array_equal(pcr.laz_decode(cut, 1500, 100), rec[1500:1600])


def test_chunk_table_pointer_at_the_end(records):
    """A writer on a stream writes -1 and appends the table's position."""
    rec, _ = expected(records, "f2_rgb_chunk1500")
    data = laz("f2_rgb_chunk1500")
    p0 = struct.unpack_from("<I", data, 96)[0]
    table = struct.unpack_from("<q", data, p0)[0]
    streamed = _patch(data, p0, "<q", -1) + struct.pack("<q", table)
    assert np.array_equal(pcr.laz_decode(streamed), rec)


def _fails(data):
    try:
        pcr.laz_decode(data)
    except pcr.PcrError:
        return True
    return False


def test_damaged_files_fail_cleanly():
    data = laz("f7_v3")
    p0 = struct.unpack_from("<I", data, 96)[0]
    # truncations: header, VLR, point data, chunk table
    for cut in (100, 300, p0 + 4, p0 + 5000, len(data) - 30, len(data) - 1):
        assert _fails(data[:cut]), cut
    # no LASzip VLR
    assert _fails(_patch(data, 100, "<I", 0))
    # an item version the decoder does not know
    hs = struct.unpack_from("<H", data, 94)[0]
    vlr = hs + 54
    assert _fails(_patch(data, vlr + 38, "<H", 1))
    # the chunk table pointer outside the file
    assert _fails(_patch(data, p0, "<q", len(data) + 100))
    # a layer size larger than the chunk
    assert _fails(_patch(data, p0 + 8 + 36 + 4, "<I", 10**8))
    with pytest.raises(pcr.PcrError, match="outside the file"):
        pcr.laz_decode(data, 4990, 20)


def test_flipped_bytes_never_crash(records):
    """Random byte flips inside chunks either fail or decode to some records."""
    rng = np.random.default_rng(5)
    for name in ("f3_extra3_variable", "f8_extra2_v3"):
        data = laz(name)
        p0 = struct.unpack_from("<I", data, 96)[0]
        table = struct.unpack_from("<q", data, p0)[0]
        for _ in range(40):
            b = bytearray(data)
            for pos in rng.integers(p0 + 8, table, 8):
                b[pos] ^= int(rng.integers(1, 256))
            try:
                out = pcr.laz_decode(bytes(b))
                assert out.shape == expected(records, name)[0].shape
            except pcr.PcrError:
                pass


@pytest.mark.skipif(not os.path.exists(REF_TOOL), reason="oracle/_ref/laz_ref_tool is built only where the "
                    "reference's sources are (make -C oracle ref)")
@pytest.mark.parametrize("name", CASES)
def test_reference_laszip_decodes_the_fixtures_alike(records, name, tmp_path):
    rec, _ = expected(records, name)
    src, out = tmp_path / "in.laz", tmp_path / "out.bin"
    src.write_bytes(laz(name))
    subprocess.run([REF_TOOL, "decompress", str(src), str(out)], check=True)
    ref = np.fromfile(out, np.uint8).reshape(rec.shape)
    assert np.array_equal(ref, rec)
    assert np.array_equal(pcr.laz_decode(laz(name)), ref)


@pytest.mark.parametrize("name,copies", [("f0_chunk1000", 3), ("f3_extra3_variable", 5), ("f7_v4", 7),
                                         ("f8_extra2_v3", 2)])
def test_tiled_file_decodes_to_repeated_records(records, name, copies):
    """datasets.laz_tile repeats a file's chunks behind a new arithmetic-coded
    chunk table (variable chunk sizes)."""
    from compute_rasterizer_amd.datasets import laz_tile
    rec, _ = expected(records, name)
    big = laz_tile(laz(name), copies)
    info = pcr.laz_info(big)
    assert info.chunk_size == 0xFFFFFFFF and info.num_points == copies * rec.shape[0]
    starts, firsts = pcr.laz_chunk_table(big)
    assert len(starts) == info.num_chunks + 1 and firsts[0] == 0 and firsts[-1] == info.num_points
    assert np.all(np.diff(starts.astype(np.int64)) > 0) and np.all(np.diff(firsts.astype(np.int64)) > 0)
    assert np.array_equal(pcr.laz_decode(big, threads=4), np.tile(rec, (copies, 1)))


def test_chunk_table_of_a_fixture():
    data = laz("f2_rgb_chunk1500")
    starts, firsts = pcr.laz_chunk_table(data)
    assert list(firsts) == [0, 1500, 3000, 4100]
    p0 = struct.unpack_from("<I", data, 96)[0]
    table = struct.unpack_from("<q", data, p0)[0]
    assert starts[0] == p0 + 8 and starts[-1] == table


@pytest.mark.skipif(not os.path.exists(REF_TOOL), reason="oracle/_ref/laz_ref_tool is built only where the "
                    "reference's sources are (make -C oracle ref)")
def test_reference_laszip_reads_a_tiled_file(records, tmp_path):
    from compute_rasterizer_amd.datasets import laz_tile
    rec, _ = expected(records, "f7_v4")
    src, out = tmp_path / "tiled.laz", tmp_path / "out.bin"
    src.write_bytes(laz_tile(laz("f7_v4"), 3))
    subprocess.run([REF_TOOL, "decompress", str(src), str(out)], check=True)
    assert np.array_equal(np.fromfile(out, np.uint8).reshape(-1, rec.shape[1]), np.tile(rec, (3, 1)))
