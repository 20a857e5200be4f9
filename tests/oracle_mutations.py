"""Mutation check of the oracle's pins (test infrastructure, run by hand).

Plants each mutation below, one at a time, in a copy of oracle/pcr_oracle.c
(in /tmp/mut_repo), rebuilds the oracle there and runs the CPU tests except
the oracle-written golden snapshot (tests/test_golden.py). Every mutation
must make a test fail. Usage: python tests/oracle_mutations.py"""
import os, shutil, subprocess, sys
SRC = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MUT = [
 ("a7 index0 bg", "(fb[i] == PCR_FB_CLEAR || id >= npoints) ? bg : rgba[id]", "(fb[i] == PCR_FB_CLEAR || id >= npoints || id == 0) ? bg : rgba[id]"),
 ("a7 force alpha", "(fb[i] == PCR_FB_CLEAR || id >= npoints) ? bg : rgba[id]", "(fb[i] == PCR_FB_CLEAR || id >= npoints) ? bg : (rgba[id] | 0xFF000000u)"),
 ("a11 rounding", "uint32_t r = acc[4 * i] / c,", "uint32_t r = (acc[4 * i] + c / 2) / c,"),
 ("a11 alpha", "out[i] = r | (g << 8) | (bb << 16);", "out[i] = r | (g << 8) | (bb << 16) | 0xFF000000u;"),
 ("a4 500->400", "if (px < 500.0f) return 2;", "if (px < 400.0f) return 2;"),
 ("a4 200->250", "if (px < 200.0f) return 3;", "if (px < 250.0f) return 3;"),
 ("a4 10000->9000", "if (px < 10000.0f) return 1;", "if (px < 9000.0f) return 1;"),
 ("a4 100->110", "orc_precision_level(b[k].min, b[k].max, v, 100.0f);\n        if (levels_out)", "orc_precision_level(b[k].min, b[k].max, v, 110.0f);\n        if (levels_out)"),
 ("a4 first<= ", "if (px < first_threshold) return 4;", "if (px <= first_threshold * 1.005f) return 4;"),
 ("a12 size80->70", "orc_precision_level(n->min, n->max, v, 80.0f) < 4;", "orc_precision_level(n->min, n->max, v, 70.0f) < 4;"),
 ("a5 level1 |512", "uint32_t X = (l << 20) | (m << 10);", "uint32_t X = (l << 20) | (m << 10) | 512u;"),
 ("a5 level0 drop lsb", "uint32_t X = (l << 20) | (m << 10) | hh;", "uint32_t X = (l << 20) | (m << 10) | (hh & ~1u);"),
 ("a5 level1 uses hi", "uint32_t X = (l << 20) | (m << 10);", "uint32_t X = (l << 20) | (m << 10) | hh;"),
 ("a12 spacing >=1.5", "v->width, v->height) > lod_px;", "v->width, v->height) >= 1.5f * lod_px;"),
 ("a12 ancestor only parent", "p = nodes[p].parent;", "p = -1;"),
 ("LAS colour <255", "c[k] = v < 256 ? v : v / 256;", "c[k] = v < 255 ? v : v / 256;"),
 ("C1 colour round", "rgba[q] = (uint32_t)(uint8_t)(255.0f * x)", "rgba[q] = (uint32_t)(uint8_t)(255.0f * x + 0.5f)"),
]
def run(name, old, new):
    d = "/tmp/mut_repo"
    if os.path.exists(d): shutil.rmtree(d)
    shutil.copytree(SRC, d, ignore=shutil.ignore_patterns(".git", "_work", "gpurun_out", "__pycache__", "_build", "_ref"))
    p = os.path.join(d, "oracle", "pcr_oracle.c")
    s = open(p).read()
    assert s.count(old) >= 1, name
    s = s.replace(old, new, 1)
    open(p, "w").write(s)
    subprocess.run(["make", "-s", "-B", "-C", os.path.join(d, "oracle")], check=True)
    r = subprocess.run([sys.executable, "-m", "pytest", "tests", "-q", "-x", "-m", "not gpu", "-p", "no:cacheprovider",
                        "--ignore=tests/test_golden.py", "--deselect", "tests/test_oracle_sanitize.py"], cwd=d,
                       capture_output=True, text=True)
    last = [l for l in r.stdout.splitlines() if l.strip()][-1]
    failed = [l for l in r.stdout.splitlines() if l.startswith("FAILED")]
    print(f"{name:28s} rc={r.returncode} {last[:60]} {failed[:1]}", flush=True)
    return r.returncode != 0
if __name__ == "__main__":
    ok = all([run(*m) for m in MUT])
    print("ALL CAUGHT" if ok else "SOME MUTATION SURVIVED")
    sys.exit(0 if ok else 1)
