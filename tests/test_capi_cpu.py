"""C ABI checks that need no GPU: every symbol declared in include/pcr.h is
exported by libpcr.so, host-side helpers agree with the oracle, and calls
without a device fail with a status code and a message."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

import compute_rasterizer_amd as pcr
from compute_rasterizer_amd import capi
from oracle import orc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "pcr.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pcr_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    out = subprocess.run(["nm", "-D", "--defined-only", capi.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r"\sT\s(pcr_\w+)", out))
    missing = [s for s in declared_symbols() if s not in exported]
    assert not missing, missing
    assert len(declared_symbols()) >= 40


def test_python_signatures_cover_header():
    assert set(declared_symbols()) <= set(capi.SIGNATURES)


# ctypes mirror -> C type in include/pcr.h
MIRRORS = {"Batch": "pcr_batch", "View": "pcr_view", "RenderParams": "pcr_render_params", "Stats": "pcr_stats",
           "GenCfg": "pcr_gen_cfg", "BuildOpts": "pcr_build_opts", "CloudInfo": "pcr_cloud_info",
           "LasInfo": "pcr_las_info", "LasFile": "pcr_las_file", "PotreeInfo": "pcr_potree_info",
           "OctreeOpts": "pcr_octree_opts", "Tune": "pcr_tune", "Request": "pcr_request",
           "ServerStats": "pcr_server_stats", "MultiViewStats": "pcr_multiview_stats",
           "LazInfo": "pcr_laz_info"}


def test_struct_layouts_match_header(tmp_path):
    """Every ctypes mirror has the size and field offsets gcc gives the header's struct."""
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "pcr.h"', "int main(void) {"]
    for py, c in MIRRORS.items():
        lines.append(f'printf("{py} __sizeof__ %zu\\n", sizeof({c}));')
        for name, _ in getattr(capi, py)._fields_:
            lines.append(f'printf("{py} {name} %zu\\n", offsetof({c}, {name}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    got = {tuple(l.split()[:2]): int(l.split()[2]) for l in out if l}
    for py in MIRRORS:
        cls = getattr(capi, py)
        assert got[(py, "__sizeof__")] == C.sizeof(cls), py
        for name, _ in cls._fields_:
            assert got[(py, name)] == getattr(cls, name).offset, (py, name)
    assert C.sizeof(capi.Batch) == 64


def test_version_and_device_count():
    lib = capi.load()
    assert lib.pcr_version() == 100
    n = C.c_int(-1)
    assert lib.pcr_device_count(C.byref(n)) == 0 and n.value >= 0


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is present")
def test_create_without_device_fails_cleanly():
    lib = capi.load()
    h = C.c_void_p()
    assert lib.pcr_create(0, C.byref(h)) == capi.PCR_E_NO_DEVICE
    assert b"no HIP device" in lib.pcr_last_error()
    with pytest.raises(pcr.PcrError):
        pcr.Context(0)


def test_null_arguments_are_rejected():
    lib = capi.load()
    assert lib.pcr_render(None, None, None, None, None) == capi.PCR_E_INVALID
    assert b"null" in lib.pcr_last_error()
    assert lib.pcr_cloud_upload_batched(None, None, 0, None, None, None, None, 0, None) == capi.PCR_E_INVALID
    assert lib.pcr_encode_batches_host(None, 0, 10240, None, None, None, None) == capi.PCR_E_INVALID


def test_view_look_at_matches_oracle_camera():
    for eye, tgt, up, w, h in (((0.5, 0.5, 2.5), (0.5, 0.5, 0.5), (0, 1, 0), 1280, 720),
                               ((1000, -200, 400), (1000, 1000, 50), (0, 0, 1), 1920, 1080),
                               ((3, 4, 5), (-1, 2, 0.5), (0, 0, 1), 257, 129)):
        a = pcr.look_at(eye, tgt, up=up, near=0.5, far=3000, width=w, height=h)
        b = orc.look_at(eye, tgt, up, near=0.5, far=3000, width=w, height=h)
        for f in ("transform", "transform_frustum", "view_world", "proj"):
            assert np.array_equal(np.array(getattr(a, f)), np.array(getattr(b, f))), f
        assert (a.width, a.height) == (w, h)


def test_view_look_at_rejects_degenerate_cameras():
    with pytest.raises(pcr.PcrError):
        pcr.look_at((0, 0, 0), (0, 0, 0))
    with pytest.raises(pcr.PcrError):
        pcr.look_at((0, 0, 0), (0, 1, 0), up=(0, 1, 0))
    with pytest.raises(pcr.PcrError):
        pcr.look_at((0, 0, 1), (0, 0, 0), near=0.0)


@pytest.mark.parametrize("n,bs", [(1, 4), (10240, 10240), (25001, 10240), (5000, 512)])
def test_host_encoder_matches_oracle(n, bs):
    cfg = pcr.gen_cfg(pcr.SCENE_TERRAIN, n, seed=n)
    xyz, _ = orc.generate(cfg)
    got = pcr.encode_batches_host(xyz, bs)
    ref = orc.encode_batches(xyz, bs)
    assert np.array_equal(got[0].view(np.uint8), ref[0].view(np.uint8))
    for g, r in zip(got[1:], ref[1:]):
        assert np.array_equal(g, r)


def test_host_encoder_rejects_unaligned_batch_size():
    with pytest.raises(pcr.PcrError):
        pcr.encode_batches_host(np.zeros((8, 3), np.float32), 10)


def test_ground_height_matches_oracle_generator():
    """The host ground height is the generator's heightfield: equal to the
    oracle's restatement bit for bit, and every generated terrain point lies
    within its jitter of it."""
    cfg = pcr.gen_cfg(pcr.SCENE_CITY, 10_000, seed=42, extent=2000.0, amplitude=150.0, jitter=0.5)
    rng = np.random.default_rng(0)
    for x, y in rng.uniform(0, 2000, (200, 2)):
        want = orc.load().orc_terrain_height(42, 2000.0, 150.0, float(np.float32(x)), float(np.float32(y)))
        assert pcr.ground_height(cfg, float(np.float32(x)), float(np.float32(y))) == want
    tcfg = pcr.gen_cfg(pcr.SCENE_TERRAIN, 5000, seed=3, jitter=0.5)
    xyz, _ = orc.generate(tcfg)
    g = np.array([pcr.ground_height(tcfg, float(p[0]), float(p[1])) for p in xyz[:500]])
    assert np.all(np.abs(xyz[:500, 2] - g) <= 1.0 + 1e-3)
    with pytest.raises(pcr.PcrError):
        pcr.ground_height(pcr.gen_cfg(pcr.SCENE_CUBE, 10), 0.0, 0.0)
