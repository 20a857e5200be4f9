"""GPU: the C++ Method adapter (include/pcr_method.hpp) driven through the
reference-style Runtime registry; each method's image equals a direct
C-ABI render of the same view (each eye's, for the two-eye VR methods)."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "method_main")


def test_cpp_methods_render_through_runtime():
    assert os.path.exists(EXE), "build() compiles tests/cpp/method_main"
    out = subprocess.run([EXE], capture_output=True, text=True, timeout=240)
    print(out.stdout, out.stderr)
    assert out.returncode == 0, out.stdout + out.stderr
    names = [l.split()[1] for l in out.stdout.splitlines() if l.startswith("OK ")]
    single = ["loop_las", "loop_las_prefetch", "loop_las_hqs", "loop_nodes", "loop_nodes_hqs"]
    # the two-eye methods print one line per frame and eye (2 frames x 2 eyes)
    assert names == single + ["loop_las_hqs_vr"] * 4 + ["loop_nodes_hqs_vr"] * 4


def test_cpp_las_files_through_loop_las_prefetch(tmp_path):
    """LasFiles reads three LAS files into one cloud; loop_las_prefetch draws
    each file with its own origin, as a direct per-file render does."""
    from compute_rasterizer_amd.datasets import make_las
    paths = []
    for k, (n, fmt, ver) in enumerate([(25_601, 2, (1, 2)), (13_333, 3, (1, 3)), (20_482, 7, (1, 4))]):
        data = make_las(n, point_format=fmt, version=ver, seed=10 + k, offset=(500000.0 + 40 * k, 4000000.0, 100.0))[0]
        p = tmp_path / f"f{k}.las"
        p.write_bytes(data)
        paths.append(str(p))
    out = subprocess.run([EXE, "--las", *paths], capture_output=True, text=True, timeout=240)
    print(out.stdout, out.stderr)
    assert out.returncode == 0 and "OK las_files 3" in out.stdout, out.stdout + out.stderr
    # a file without points, in front: no part, and the same image
    empty = tmp_path / "empty.las"
    empty.write_bytes(make_las(0, point_format=2, version=(1, 2), seed=9)[0])
    out2 = subprocess.run([EXE, "--las", str(empty), *paths], capture_output=True, text=True, timeout=240)
    print(out2.stdout, out2.stderr)
    assert out2.returncode == 0 and "OK las_files 3" in out2.stdout, out2.stdout + out2.stderr
    assert out.stdout.split()[-1] == out2.stdout.split()[-1]


def test_cpp_potree_through_loop_nodes(tmp_path):
    """PotreeCloud reads a Potree 2.0 dataset; loop_nodes renders it as a
    direct render of the same cloud does."""
    import compute_rasterizer_amd as pcr
    from compute_rasterizer_amd.datasets import make_potree
    meta, hier, octree, nodes = make_potree(seed=12, max_depth=5)
    (tmp_path / "hierarchy.bin").write_bytes(hier)
    (tmp_path / "octree.bin").write_bytes(octree)
    (tmp_path / "info.bin").write_bytes(bytes(pcr.potree_info(meta)))
    out = subprocess.run([EXE, "--potree", str(tmp_path), str(tmp_path / "info.bin")], capture_output=True,
                         text=True, timeout=240)
    print(out.stdout, out.stderr)
    assert out.returncode == 0 and "OK potree" in out.stdout, out.stdout + out.stderr
