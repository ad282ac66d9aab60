"""Checks of the B-spline order 5 builds (libmpid_hip_order5.so and the
oracle's libmpid_oracle_order5.so, a compile-time option; the reference and
the default builds use order 6).  The library and the oracle pick their
build at import from MPID_PME_ORDER, so these checks run in a process of
their own: tests/test_oracle_order5.py and tests/test_gpu_order5.py start it
with MPID_PME_ORDER=5.  Asserts, and prints one JSON line of the measured
differences.

    python tests/order5_worker.py oracle   # the oracle's order-5 pins (CPU)
    python tests/order5_worker.py gpu      # the GPU build against the oracle
"""
import json
import os
import sys
from math import comb, factorial

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
assert os.environ.get("MPID_PME_ORDER") == "5", "run with MPID_PME_ORDER=5"

from oracle import oracle  # noqa: E402
from mpidopenmmplugin_amd.waterbox import make_waterbox  # noqa: E402


def rel_rms(a, b):
    return float(np.sqrt(((np.asarray(a) - np.asarray(b)) ** 2).sum() / (np.asarray(b) ** 2).sum()))


def m5(x):
    """The cardinal B-spline of order 5 (degree 4), closed form."""
    return sum((-1) ** j * comb(5, j) * max(x - j, 0.0) ** 4 for j in range(6)) / factorial(4)


def oracle_checks():
    out = {}
    assert oracle.PME_ORDER == 5
    for w in (0.0, 0.13, 0.5, 0.77, 0.999):
        th = oracle.bspline(w)
        assert th.shape == (5, 5)
        assert abs(th[:, 0].sum() - 1.0) < 1e-14
        for r in range(1, 5):
            assert abs(th[:, r].sum()) < 1e-13
        for k in range(5):
            assert abs(th[k, 0] - m5(w + 4 - k)) < 1e-13, (w, k)
        h = 1e-6
        if 1e-5 < w < 1 - 1e-5:
            thp, thm = oracle.bspline(w + h), oracle.bspline(w - h)
            for r in range(1, 4):  # the 4th derivative of a quartic piece is constant: its FD is the 3rd's
                fd = (thp[:, r - 1] - thm[:, r - 1]) / (2 * h)
                assert np.allclose(fd, th[:, r], atol=1e-6), (w, r)
    # at w = 0 the weights are M5 at the integers: 1/24, 11/24, 11/24, 1/24
    th0 = oracle.bspline(0.0)[:, 0]
    assert np.allclose(sorted(th0)[1:], sorted([1 / 24, 11 / 24, 11 / 24, 1 / 24]), atol=1e-15)
    # order-5 moduli vanish at the Nyquist frequency of an even grid; the
    # reference's fix (half the neighbours' sum) keeps them positive
    m = oracle.bspline_moduli(24)
    assert m.min() > 0.0
    out["moduli_min_24"] = float(m.min())
    # energies on refined grids (the parent compares them with order 6)
    f, pos, box = make_waterbox("water216")
    for g in (24, 48):
        f.setPMEParameters(f.getPMEParameters()[0], g, g, g)
        out[f"energy_{g}"] = oracle.evaluate(f.flatten(), pos, box)["energy"]
    return out


def gpu_checks():
    from mpidopenmmplugin_amd import _capi
    from mpidopenmmplugin_amd.kernel import HipCalcMPIDForceKernel

    assert _capi.load().mpid_get_pme_bspline_order() == 5
    out = {}

    def run(force, pos, box, precision, tiling="library"):
        k = HipCalcMPIDForceKernel(precision, spread_tiling=tiling)
        k.initialize(force.getNumMultipoles(), force, default_box=box)
        e, F = k.execute(pos, box)
        return e, F, k.getInducedDipoles(pos, box)

    # water216, 24^3: FP64 converged like the oracle (1e-10 Debye), every tiling
    f, pos, box = make_waterbox("water216", target_eps=1e-10)
    ref = oracle.evaluate(f.flatten(), pos, box, details=True)
    grid = tuple(f.flatten()["grid"])
    first = None
    for t in _capi.spread_tilings(grid):
        e, F, mu = run(f, pos, box, "double", t)
        de, dF, dmu = abs(e - ref["energy"]) / abs(ref["energy"]), rel_rms(F, ref["forces"]), rel_rms(mu, ref["induced"])
        out[f"water216_double_{t[0]}_{t[1]}"] = [de, dF, dmu]
        assert de <= 1e-9 and dF <= 1e-8 and dmu <= 1e-8, (t, de, dF, dmu)
        if first is None:
            first = (e, F)
        else:
            assert abs(e - first[0]) <= 1e-12 * abs(first[0]) and rel_rms(F, first[1]) <= 1e-10, t
    # mixed at the workload's target against the same tight oracle
    g, _, _ = make_waterbox("water216")
    e, F, mu = run(g, pos, box, "mixed")
    de, dF, dmu = abs(e - ref["energy"]) / abs(ref["energy"]), rel_rms(F, ref["forces"]), rel_rms(mu, ref["induced"])
    out["water216_mixed"] = [de, dF, dmu]
    assert de <= 1e-6 and dF <= 1e-5 and dmu <= 1e-5, (de, dF, dmu)
    # water4096, 64^3 (owned cores of 8^3 points) at the workload's target
    f4, pos4, box4 = make_waterbox("water4096")
    ref4 = oracle.evaluate(f4.flatten(), pos4, box4, details=True)
    for precision in ("double", "mixed"):
        e, F, mu = run(f4, pos4, box4, precision, None)
        de, dF, dmu = abs(e - ref4["energy"]) / abs(ref4["energy"]), rel_rms(F, ref4["forces"]), rel_rms(mu, ref4["induced"])
        out[f"water4096_{precision}"] = [de, dF, dmu]
        assert de <= 1e-6 and dF <= 1e-5 and dmu <= 1e-5, (precision, de, dF, dmu)
    return out


if __name__ == "__main__":
    res = oracle_checks() if sys.argv[1] == "oracle" else gpu_checks()
    print(json.dumps(res), flush=True)
