"""The B-spline order 5 build on the GPU (libmpid_hip_order5.so, a
compile-time option; the reference hard-wires order 6) against the oracle's
order-5 build: water216 (24^3) in FP64 converged to 1e-10 Debye for every
spreading tiling (energy within 1e-9 relative, forces and induced dipoles
within 1e-8 relative RMS; the tilings agree with each other to 1e-12 in
energy), mixed precision at 1e-6 Debye (energy 1e-6, forces and dipoles
1e-5), and water4096 (64^3, owned cores) in both precisions at 1e-6 Debye
(energy 1e-6, forces and dipoles 1e-5).  The checks run in
tests/order5_worker.py, in a process of its own (the library and the oracle
pick their build at import)."""
import json
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_order5_build_matches_order5_oracle():
    print("[gpu] order-5 build (subprocess): water216 every tiling, water4096", flush=True)
    env = dict(os.environ, MPID_PME_ORDER="5")
    p = subprocess.run([sys.executable, "-u", os.path.join(ROOT, "tests", "order5_worker.py"), "gpu"], env=env,
                       capture_output=True, text=True, timeout=500)
    print(p.stdout[-2000:], p.stderr[-2000:])
    assert p.returncode == 0, p.stdout + p.stderr
    res = json.loads(p.stdout.strip().splitlines()[-1])
    assert "water4096_mixed" in res and "water216_mixed" in res
