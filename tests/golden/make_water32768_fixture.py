"""Writes tests/golden/water32768_oracle.npz: the FP64 oracle's energy,
forces and induced dipoles for the 98,304-atom water box (9.932 nm, 128^3
grid, 0.9 nm cutoff, octopoles, anisotropic polarisability), mutual
polarisation converged to 1e-10 Debye.  The oracle needs several minutes
for it, far longer than a GPU test should spend, so the result is committed
as a fixture.  The permanent field is left out to keep the file small.

Run:  python tests/golden/make_water32768_fixture.py
"""
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from mpidopenmmplugin_amd.waterbox import make_waterbox  # noqa: E402
from oracle import oracle  # noqa: E402


def main():
    f, pos, box = make_waterbox("water32768", target_eps=1e-10)
    t = time.time()
    res = oracle.evaluate(f.flatten(), pos, box, details=True)
    print(f"oracle: {time.time() - t:.1f} s, {res['iterations']} DIIS iterations, eps {res['eps']:.2e}")
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "water32768_oracle.npz")
    np.savez_compressed(out, energy=np.float64(res["energy"]), energy_terms=np.array(res["energy_terms"]),
                        forces=res["forces"], induced=res["induced"], positions_checksum=np.float64(pos.sum()))
    print("wrote", out, os.path.getsize(out), "bytes")


if __name__ == "__main__":
    main()
