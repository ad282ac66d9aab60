"""The per-pair operation counts that bench.py's roofline uses are the ones
the operation counter (csrc/opcount.cpp) measures on the shipped pair
formulas, and they respond to the formulas' branches as they should."""
import os
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.join(os.path.dirname(__file__), "..")
CSRC = os.path.join(ROOT, "mpidopenmmplugin_amd", "csrc")


@pytest.fixture(scope="module")
def counts():
    if shutil.which("hipcc") is None and not os.path.exists("/opt/rocm/bin/hipcc"):
        pytest.skip("hipcc not available")
    out = subprocess.run(["make", "-s", "-C", CSRC, "opcount"], check=True, capture_output=True, text=True).stdout
    res = {}
    for line in out.splitlines():
        parts = line.split()
        if len(parts) == 3 and parts[0].endswith("_pair"):
            res[parts[0]] = (int(parts[1]), int(parts[2]))
    return res


def test_bench_uses_the_measured_counts(counts):
    sys.path.insert(0, ROOT)
    import bench
    assert bench.PAIR_FLOPS["fixed_pairs"] == counts["fixed_field_pair"]
    assert bench.PAIR_FLOPS["induced_pairs"] == counts["induced_field_pair"]
    assert bench.PAIR_FLOPS["electrostatic_pairs"] == counts["elec_pair"]


def test_counts_follow_the_thole_branch(counts):
    # damping terms cost operations only when both sites are polarisable
    for name in ("fixed_field_pair", "induced_field_pair"):
        damped, undamped = counts[name]
        assert damped > undamped > 0
    # the electrostatics clamps the Thole argument instead of branching: one exp apart
    d, u = counts["elec_pair"]
    assert 0 < d - u <= 4
