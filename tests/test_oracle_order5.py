"""The oracle's B-spline order 5 build (a compile-time option; the reference
hard-wires order 6), pinned by what the mathematics fixes: the order-5
cardinal B-spline in closed form and its derivatives, positive moduli after
the Nyquist fix, and an energy that approaches the order-6 one as the grid
is refined (both approximate the same Ewald sum).  Runs tests/order5_worker.py
in a process of its own (the build is chosen at import)."""
import json
import os
import subprocess
import sys

from oracle import oracle
from mpidopenmmplugin_amd.waterbox import make_waterbox

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_order5_oracle():
    env = dict(os.environ, MPID_PME_ORDER="5")
    p = subprocess.run([sys.executable, os.path.join(ROOT, "tests", "order5_worker.py"), "oracle"], env=env,
                       capture_output=True, text=True, timeout=600)
    assert p.returncode == 0, p.stdout + p.stderr
    res = json.loads(p.stdout.strip().splitlines()[-1])
    assert oracle.PME_ORDER == 6
    f, pos, box = make_waterbox("water216")
    d = {}
    for g in (24, 48):
        f.setPMEParameters(f.getPMEParameters()[0], g, g, g)
        e6 = oracle.evaluate(f.flatten(), pos, box)["energy"]
        d[g] = abs(res[f"energy_{g}"] - e6) / abs(e6)
    # the orders differ on the coarse grid and agree better on the fine one
    assert d[24] > 1e-9
    assert d[48] < d[24] and d[48] < 1e-5, d
