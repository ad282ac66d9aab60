"""The tilings of the PME spreading on the GPU.

Every tiling that mpid_spread_tilings offers for a grid (owned cores,
windows folded into the grid, or atomic adds without tiles) must give the
same energy, forces and induced dipoles: in FP64 they differ only in the
order of the grid sums (energy within 1e-12 relative, forces and dipoles
within 1e-10 relative RMS).  The batched workload's largest systems (1,024
and 798 waters on 40^3 and 36^3 grids, spread by owned cores) are checked
against the FP64 oracle at the batch tests' tolerances: energy within 1e-6
relative, forces and induced dipoles within 1e-5 relative RMS of the oracle
converged to 1e-9 Debye.
"""
import numpy as np
import pytest

from mpidopenmmplugin_amd import _capi
from mpidopenmmplugin_amd.kernel import HipCalcMPIDForceKernel
from mpidopenmmplugin_amd.waterbox import make_small_waterbox
from oracle import oracle

pytestmark = pytest.mark.gpu

# (waters, seed): PME grids 40^3, 36^3, 24^3, 20^3
CASES = [(1024, 11), (798, 12), (200, 13), (120, 14)]


def rel_rms(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return np.sqrt(((a - b) ** 2).sum() / (b ** 2).sum())


def run(force, pos, box, precision, tiling):
    k = HipCalcMPIDForceKernel(precision, spread_tiling=tiling)
    k.initialize(force.getNumMultipoles(), force, default_box=box)
    e, F = k.execute(pos, box)
    mu = k.getInducedDipoles(pos, box)
    return e, F, mu


@pytest.mark.parametrize("nw,seed", CASES)
def test_every_tiling_gives_the_same_result(nw, seed):
    force, pos, box = make_small_waterbox(nw, seed)
    grid = force.flatten()["grid"]
    tilings = _capi.spread_tilings(grid)
    print(f"[gpu] tilings {nw} waters, grid {tuple(grid)}: {tilings}", flush=True)
    assert len(tilings) >= 2
    e0, F0, mu0 = run(force, pos, box, "double", tilings[0])
    for t in tilings[1:]:
        e, F, mu = run(force, pos, box, "double", t)
        assert abs(e - e0) <= 1e-12 * abs(e0), (t, e, e0)
        assert rel_rms(F, F0) <= 1e-10, t
        assert rel_rms(mu, mu0) <= 1e-10, t


@pytest.fixture(scope="module")
def large_refs():
    out = []
    for nw, seed in CASES[:2]:
        f, pos, box = make_small_waterbox(nw, seed)
        tight, _, _ = make_small_waterbox(nw, seed, target_eps=1e-9)
        tight.setMutualInducedMaxIterations(300)
        out.append((nw, f, pos, box, oracle.evaluate(tight.flatten(), pos, box, details=True)))
    return out


@pytest.mark.parametrize("precision", ["double", "mixed"])
def test_largest_batch_systems_match_oracle(large_refs, precision):
    for nw, f, pos, box, ref in large_refs:
        grid = tuple(f.flatten()["grid"])
        assert _capi.spread_tilings(grid)[0] == (2, 1), grid  # owned cores, 2^2 tiles per axis
        print(f"[gpu] oracle {nw} waters, grid {grid}, {precision}", flush=True)
        e, F, mu = run(f, pos, box, precision, "library")
        assert abs(e - ref["energy"]) <= 1e-6 * abs(ref["energy"]), (nw, e, ref["energy"])
        assert rel_rms(F, ref["forces"]) <= 1e-5, nw
        assert rel_rms(mu, ref["induced"]) <= 1e-5, nw


def test_tiling_not_offered_is_refused():
    force, pos, box = make_small_waterbox(120, 14)  # 20^3: no windows at level 1
    grid = force.flatten()["grid"]
    assert (1, 2) not in _capi.spread_tilings(grid)
    k = HipCalcMPIDForceKernel("double", spread_tiling=(1, 2))
    with pytest.raises(Exception, match="not available"):
        k.initialize(force.getNumMultipoles(), force, default_box=box)
