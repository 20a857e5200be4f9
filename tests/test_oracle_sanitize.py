"""The CPU oracle under AddressSanitizer + UndefinedBehaviorSanitizer (gcc, CPU
only): tests/cpp/oracle_sanitize.c drives every oracle entry point, including
its edge cases, and must exit cleanly with no sanitizer report."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not available")
def test_oracle_clean_under_asan_ubsan(tmp_path):
    exe = str(tmp_path / "oracle_sanitize")
    subprocess.run(["gcc", "-std=c11", "-O1", "-g", "-ffp-contract=off", "-mfma", "-fno-omit-frame-pointer",
                    "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
                    os.path.join(ROOT, "tests", "cpp", "oracle_sanitize.c"),
                    os.path.join(ROOT, "oracle", "pcr_oracle.c"), "-lm", "-lpthread", "-o", exe], check=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:verify_asan_link_order=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    out = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "oracle sanitize ok" in out.stdout
    assert "runtime error" not in out.stderr
