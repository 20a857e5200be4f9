"""GPU: the C++ Method adapter (include/pcr_method.hpp) driven through the
reference-style Runtime registry; each method's image equals a direct
C-ABI render of the same view."""
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
    assert names == ["loop_las", "loop_las_prefetch", "loop_las_hqs", "loop_nodes"]
