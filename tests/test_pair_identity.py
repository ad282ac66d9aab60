"""The reduced electrostatics formula for a pair with one bare charge
(elec_pair_bare in csrc/pair_math.h) is the general quasi-internal-frame
formula (elec_pair) with the bare site's higher moments and induced dipole
set to zero, and the bare site takes no torque.  Checked by a host program
(csrc/pair_identity.cpp) that instantiates both in FP64 over random pairs,
scale factors, Thole widths and both site orders."""
import os
import shutil
import subprocess

import pytest

CSRC = os.path.join(os.path.dirname(__file__), "..", "mpidopenmmplugin_amd", "csrc")


def test_bare_charge_formula_matches_the_general_one():
    if shutil.which("hipcc") is None and not os.path.exists("/opt/rocm/bin/hipcc"):
        pytest.skip("hipcc not available")
    out = subprocess.run(["make", "-s", "-C", CSRC, "pair_identity"], check=True, capture_output=True,
                         text=True).stdout
    v = {}
    for line in out.splitlines():
        parts = line.split()
        if len(parts) == 2:
            v[parts[0]] = float(parts[1])
    # the outputs are not trivially zero
    assert v["energy_mag"] > 1.0 and v["force_mag"] > 1.0 and v["torque_mag"] > 1.0
    # same operations in the same order for the surviving terms: agreement to rounding
    assert v["energy_rel"] < 1e-13
    assert v["force_rel"] < 1e-13
    assert v["torque_rel"] < 1e-13
    assert v["bare_torque_abs"] == 0.0
    # the Cartesian form of the tile kernel, a different order of operations
    for k in ("cart_energy_rel", "cart_force_rel", "cart_torque_rel", "cart_pot_rel"):
        assert v[k] < 1e-13, k
