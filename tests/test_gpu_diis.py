"""The DIIS solver mode (MPID_FLAG_DIIS, the reference's iteration for
mutual polarization) against the oracle's DIIS, which restates the
reference's (MPIDReferenceForce.cpp:1182-1291).  At the same target (1e-6
Debye) both follow the same iterates, so the iteration count equals the
oracle's, the final error agrees to 1e-3 of itself, and the energy (with
the same variational correction), forces and induced dipoles agree to
rounding: energy within 1e-10 relative, forces and dipoles within 1e-8
relative RMS.  Two different solvers stopped at the same target differ by
about the target instead (the CG mode agrees with the oracle to 1e-6).
Mixed precision with DIIS is refused, and the iteration limit applies."""
import numpy as np
import pytest

from mpidopenmmplugin_amd.kernel import HipCalcMPIDForceKernel, MPIDConvergenceError, MPIDUnsupportedError
from mpidopenmmplugin_amd.waterbox import make_waterbox
from oracle import oracle

pytestmark = pytest.mark.gpu


def rel_rms(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return np.sqrt(((a - b) ** 2).sum() / (b ** 2).sum())


@pytest.mark.parametrize("name", ["water216", "water4096"])
def test_diis_follows_the_oracle(name):
    f, pos, box = make_waterbox(name)
    print(f"[gpu] DIIS {name}", flush=True)
    ref = oracle.evaluate(f.flatten(), pos, box, details=True)
    k = HipCalcMPIDForceKernel("double", solver="diis")
    k.initialize(len(pos), f)
    e, F = k.execute(pos, box)
    mu = k.getInducedDipoles(pos, box)
    st = k.stats()
    assert st["iterations"] == ref["iterations"], (st["iterations"], ref["iterations"])
    # the final error is a small difference of nearly equal vectors (about
    # 3e-7 of the dipoles), and the two operators agree to about 1e-10: the
    # error agrees to about 1e-4 of itself
    assert abs(st["epsilon"] - ref["eps"]) <= 1e-3 * ref["eps"], (st["epsilon"], ref["eps"])
    assert abs(e - ref["energy_variational"]) <= 1e-10 * abs(ref["energy_variational"]), (e, ref["energy_variational"])
    assert rel_rms(F, ref["forces"]) <= 1e-8
    assert rel_rms(mu, ref["induced"]) <= 1e-8


def test_diis_is_double_precision_only():
    f, pos, box = make_waterbox("water216")
    k = HipCalcMPIDForceKernel("mixed", solver="diis")
    with pytest.raises(MPIDUnsupportedError, match="DIIS"):
        k.initialize(len(pos), f)


def test_diis_iteration_limit():
    f, pos, box = make_waterbox("water216")
    ref = oracle.evaluate(f.flatten(), pos, box)
    g, _, _ = make_waterbox("water216")
    g.setMutualInducedMaxIterations(ref["iterations"] - 1)
    k = HipCalcMPIDForceKernel("double", solver="diis")
    k.initialize(len(pos), g)
    with pytest.raises(MPIDConvergenceError, match="did not converge"):
        k.execute(pos, box)
