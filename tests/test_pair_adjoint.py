"""The pair formulas behind the energy gradients (csrc/pair_math.h), checked
on the host in FP64 by csrc/pair_adjoint.cpp against what linear algebra and
calculus fix: the QI-frame potential transform is the transpose of the moment
transform; the potential vectors of the general and one-bare-charge pair
formulas are the derivatives of the pair energy with the induced dipoles
interacting fully (central differences, exact for an energy linear in each
moment); and the Thole-width derivative matches central differences of the
damped field energy."""
import os
import shutil
import subprocess

import pytest

CSRC = os.path.join(os.path.dirname(__file__), "..", "mpidopenmmplugin_amd", "csrc")


def test_gradient_pair_formulas():
    if shutil.which("hipcc") is None and not os.path.exists("/opt/rocm/bin/hipcc"):
        pytest.skip("hipcc not available")
    out = subprocess.run(["make", "-s", "-C", CSRC, "pair_adjoint"], check=True, capture_output=True,
                         text=True).stdout
    v = {}
    for line in out.splitlines():
        parts = line.split()
        if len(parts) == 2:
            v[parts[0]] = float(parts[1])
    assert v["qi_transpose_rel"] < 1e-13
    # not trivially zero
    assert v["pot_general_mag"] > 1.0 and v["pot_bare_mag"] > 1.0 and v["thole_mag"] > 1e-3
    assert v["pot_general_rel"] < 1e-12
    assert v["pot_bare_rel"] < 1e-12
    # second-order central differences with a 1e-5 relative step
    assert v["thole_rel"] < 1e-8
