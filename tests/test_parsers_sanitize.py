"""The library's host parsers under AddressSanitizer + UndefinedBehaviorSanitizer
(g++, CPU only): the LAS header parser, the Potree 2.0 hierarchy parser and
the LAZ decoder (compute_rasterizer_amd/csrc/pcr_parse.cpp, pcr_laz.cpp) are
built with tests/cpp/parsers_sanitize.cpp, which feeds them thousands of
mutated LAS headers, hierarchies and LAZ files. Every call must return, with
no sanitizer report."""
import ctypes as C
import os
import shutil
import subprocess

import pytest

import compute_rasterizer_amd as pcr
from compute_rasterizer_amd.datasets import make_las, make_potree

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "compute_rasterizer_amd", "csrc")
LAZ_DIR = os.path.join(ROOT, "tests", "golden", "laz")


@pytest.mark.skipif(shutil.which("g++") is None or not os.path.exists("/opt/rocm/include/hip/hip_runtime.h"),
                    reason="g++ or the HIP headers are not available")
def test_host_parsers_clean_under_asan_ubsan(tmp_path):
    exe = str(tmp_path / "parsers_sanitize")
    subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
                    "-fno-sanitize-recover=all", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", "-I", CSRC,
                    os.path.join(ROOT, "tests", "cpp", "parsers_sanitize.cpp"), os.path.join(CSRC, "pcr_parse.cpp"),
                    os.path.join(CSRC, "pcr_laz.cpp"), "-lpthread", "-o", exe], check=True)
    args = []
    for k, (fmt, ver) in enumerate(((2, (1, 2)), (3, (1, 3)), (7, (1, 4)), (8, (1, 4)))):
        p = tmp_path / f"seed{k}.las"
        p.write_bytes(make_las(64, point_format=fmt, version=ver, seed=k, extra_bytes=k)[0])
        args.append(f"las:{p}")
    for k, (cd, md) in enumerate(((2, 4), (1, 3))):
        meta, hier, octree, _ = make_potree(seed=20 + k, chunk_depth=cd, max_depth=md)
        info = pcr.potree_info(meta)
        ip, hp = tmp_path / f"info{k}.bin", tmp_path / f"hier{k}.bin"
        ip.write_bytes(C.string_at(C.addressof(info), C.sizeof(info)))
        hp.write_bytes(hier)
        args.append(f"potree:{ip}:{hp}:{len(octree)}")
    for name in ("f0_chunk1000", "f1_gps", "f3_extra3_variable", "f6_channels", "f7_v3", "f8_extra2_v4"):
        args.append(f"laz:{os.path.join(LAZ_DIR, name + '.laz')}")
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:verify_asan_link_order=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    out = subprocess.run([exe, "400"] + args, capture_output=True, text=True, env=env, timeout=600)
    assert out.returncode == 0, out.stdout[-2000:] + out.stderr[-4000:]
    assert "parsers sanitize ok" in out.stdout
    assert "runtime error" not in out.stderr
    print(out.stdout.strip())
