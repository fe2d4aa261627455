"""Write the pyramid golden vectors from the reference's numpy pyramid.

Run once, in the development container where the reference tree exists:

    python tests/golden/make_pyramid_golden.py /path/to/reference

It imports code/python/poc/dmapref/pyramid.py (numpy only) and writes, for
each case, the node grid, the scale and all eight channels of every level
for both builds into tests/golden/pyramid_<case>.npz. The cases follow the
reference's own exporter (code/python/poc/experiments/export_golden_pyramid.py:
39-48): random_65 is default_rng(2027).random((65, 65)) at scale 1.0, the one
case of that exporter that needs no image file. Two small extra cases cover
an adversarial step edge (the grid of code/python/poc/tests/test_pyramid.py:
85-90) and a seeded 17^2 grid at scale 0.7.

Only the h0 and r channels of the fold need these vectors: every other
channel is also pinned by invariants in tests/test_oracle_pyramid.py.
"""
import sys
from pathlib import Path

import numpy as np


def main(ref_root):
    sys.path.insert(0, str(Path(ref_root) / "code" / "python" / "poc"))
    from dmapref.pyramid import BUILDS, CHANNELS, TaylorPyramid

    step = np.zeros((17, 17))
    step[:, 9:] = 1.0
    cases = {
        "random_65": (np.random.default_rng(2027).random((65, 65)), 1.0),
        "step_17": (step, 1.0),
        "random_17": (np.random.default_rng(11).random((17, 17)), 0.7),
    }
    out_dir = Path(__file__).resolve().parent
    for name, (values, scale) in cases.items():
        arrays = {"values": np.asarray(values, dtype=np.float64), "scale": np.array([scale])}
        for build in BUILDS:
            pyr = TaylorPyramid(values, scale, build)
            for k, level in enumerate(pyr.levels):
                for ch in CHANNELS:
                    arrays[f"{build}/{ch}_L{k}"] = np.asarray(level[ch], dtype=np.float64)
        np.savez_compressed(out_dir / f"pyramid_{name}.npz", **arrays)
        print(f"wrote pyramid_{name}.npz ({len(arrays)} arrays)")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
