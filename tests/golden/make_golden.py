"""Writes tests/golden/golden_small.npz: small seeded clouds, several cameras and
the oracle's framebuffers / images for the plain (16 B/pt), batched and hqs
paths. Run from the repo root: python tests/golden/make_golden.py

The fixture is data only (numpy arrays, no pickles). The oracle output it
records is independently pinned by tests/test_oracle.py; the fixture guards
against regressions and checks that the oracle and the GPU produce the same
bits on other machines.
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from compute_rasterizer_amd.capi import GenCfg  # noqa: E402
from oracle import orc  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden_small.npz")

CAMERAS = [  # eye, target, up, width, height
    ((0.5, 0.5, 2.5), (0.5, 0.5, 0.5), (0, 1, 0), 128, 72),
    ((2.0, 1.2, 1.5), (0.5, 0.5, 0.5), (0, 1, 0), 96, 96),
    ((0.6, 0.4, 1.1), (0.5, 0.5, 0.4), (0, 1, 0), 160, 90),
]


def views():
    return [orc.look_at(e, t, u, fovy=60.0, near=0.1, far=1000.0, width=w, height=h) for e, t, u, w, h in CAMERAS]


def build():
    cfg = GenCfg(0, 0, 3, 4096, 0.0, 0.0, 0.0, 0.0)        # CUBE, seed 3, 4096 points
    xyz, rgba = orc.generate(cfg)
    b, lo, med, hi = orc.encode_batches(xyz, 1024)
    out = {"xyz": xyz, "rgba": rgba, "batches": b.view(np.uint8), "lo": lo, "med": med, "hi": hi}
    for k, v in enumerate(views()):
        out[f"view{k}"] = np.frombuffer(bytes(v), np.uint8)
        fb16 = orc.render_points16(xyz, rgba, v)
        out[f"plain_keys{k}"] = fb16
        out[f"plain_rgba{k}"] = orc.resolve_color(fb16, 0xFF000000)
        fbb, levels = orc.render_batched(b, lo, med, hi, v, 1)
        out[f"batched_keys{k}"] = fbb
        out[f"batched_levels{k}"] = levels
        out[f"batched_rgba{k}"] = orc.resolve_index(fbb, rgba, 0x00443322)
        d, a = orc.hqs_points16(xyz, rgba, v)
        out[f"hqs_depth{k}"] = d
        out[f"hqs_acc{k}"] = a
        out[f"hqs_rgba{k}"] = orc.hqs_resolve(a, 0x00443322)
    return out


if __name__ == "__main__":
    np.savez_compressed(OUT, **build())
    print("wrote", OUT)
