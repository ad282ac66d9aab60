"""Writes tests/golden/frames_gradients_fd.npz: central finite differences of
the FP64 oracle's variational energy for the frames cluster
(tests/helpers.py frames_cluster: ThreeFold, ZBisect, ZThenX with a y atom
in both handednesses, Bisector, ZOnly and NoAxisType sites, every site
polarisable), without cutoff and with PME, converged to 1e-10 Debye at every
point.  Sites, parameter classes, steps and stencils are those of
tests/test_gpu_gradients.py (FRAME_SITES, FRAME_CLASSES, STEPS, FOUR_POINT).

The positions, the box and the flattened parameters that the differences
were taken at are stored beside them ("<case>_in_<name>"), so that the test
can check that it evaluates the same system.

Run from the repository root (about eight minutes on one core):

    python tests/golden/make_frames_fd_fixture.py
"""
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from tests.helpers import frames_cluster  # noqa: E402
from tests.test_gpu_gradients import FRAME_CLASSES, FRAME_INPUTS, FRAME_SITES, fd_gradients, frame_box  # noqa: E402

OUT = os.path.join(ROOT, "tests", "golden", "frames_gradients_fd.npz")


def main():
    arrays = {}
    for periodic, case in ((False, "nocutoff"), (True, "pme")):
        t = time.perf_counter()
        force, pos, box = frames_cluster(periodic=periodic, target_eps=1e-10)
        flat = force.flatten()
        fd = fd_gradients(flat, pos, box, FRAME_SITES, FRAME_CLASSES)
        for cls, g in fd.items():
            arrays[f"{case}_{cls}"] = g
        arrays[f"{case}_in_positions"] = np.asarray(pos, dtype=np.float64)
        arrays[f"{case}_in_box"] = frame_box(box)
        for name in FRAME_INPUTS:
            arrays[f"{case}_in_{name}"] = np.asarray(flat[name], dtype=np.float64)
        print(f"{case}: {time.perf_counter() - t:.1f} s", flush=True)
    np.savez_compressed(OUT, **arrays)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
