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
        if len(parts) == 3 and parts[0].startswith(("fixed", "induced", "elec", "charge")):
            res[parts[0]] = (int(parts[1]), int(parts[2]))
    return res


def test_bench_uses_the_measured_counts(counts):
    sys.path.insert(0, ROOT)
    import bench
    assert set(bench.PAIR_FLOPS) == set(counts)
    for name, value in counts.items():
        assert bench.PAIR_FLOPS[name] == value, name


def test_counts_follow_the_thole_branch(counts):
    # damping terms cost operations only when both sites are polarisable
    for name in ("fixed_field_radial", "induced_field_radial"):
        damped, undamped = counts[name]
        assert damped > undamped > 0
    for name in ("fixed_field_from", "induced_field_from", "charge_pair"):
        assert counts[name][0] == counts[name][1] > 0
    # the electrostatics clamps the Thole argument instead of branching: one exp apart
    d, u = counts["elec_pair"]
    assert 0 < d - u <= 4
    # a pair of bare charges costs a small fraction of the general interaction
    assert counts["charge_pair"][0] * 50 < counts["elec_pair"][1]


def test_cartesian_bare_formula_is_cheaper(counts):
    assert 0 < counts["elec_pair_bare_cart"][0] * 2 < counts["elec_pair_bare"][0]
