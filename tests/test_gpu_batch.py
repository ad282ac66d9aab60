"""The batched workload on the GPU: small periodic water boxes (cutoffs just
under half the box, odd 2,3,5-smooth PME grids such as 15^3 and 25^3 that
take the untiled spreading path) against the FP64 oracle, and the thread-pool
evaluator against each system's own context.

Tolerances: at the workload's convergence target (1e-6 Debye) energy within
1e-6 relative, forces and induced dipoles within 1e-5 relative RMS of the
oracle converged to 1e-9 Debye, for both precisions.
"""
import numpy as np
import pytest

from mpidopenmmplugin_amd.batch import BatchEvaluator
from mpidopenmmplugin_amd.kernel import MPIDContext
from mpidopenmmplugin_amd.waterbox import make_batch, make_small_waterbox
from oracle import oracle

pytestmark = pytest.mark.gpu

# (waters, seed): grids 15, 15, 20, 25, 30 -> untiled (odd) and tiled spreading
SMALL = [(32, 1), (50, 2), (100, 5), (250, 3), (400, 4)]


def rel_rms(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return np.sqrt(((a - b) ** 2).sum() / (b ** 2).sum())


@pytest.fixture(scope="module")
def small_systems():
    systems, refs = [], []
    for nw, seed in SMALL:
        f, pos, box = make_small_waterbox(nw, seed)
        tight, _, _ = make_small_waterbox(nw, seed, target_eps=1e-9)
        tight.setMutualInducedMaxIterations(300)
        systems.append((f, pos, box))
        refs.append(oracle.evaluate(tight.flatten(), pos, box, details=True))
    return systems, refs


@pytest.mark.parametrize("precision", ["double", "mixed"])
def test_small_boxes_match_oracle(small_systems, precision):
    systems, refs = small_systems
    print(f"[gpu] batch small boxes {precision}: {[nw for nw, _ in SMALL]} waters", flush=True)
    with BatchEvaluator(systems, precision=precision, threads=3) as ev:
        res = ev.execute()
        mus = ev.induced_dipoles()
    for (nw, _), (e, F), mu, ref in zip(SMALL, res, mus, refs):
        assert abs(e - ref["energy"]) <= 1e-6 * abs(ref["energy"]), (nw, e, ref["energy"])
        assert rel_rms(F, ref["forces"]) <= 1e-5, nw
        assert rel_rms(mu, ref["induced"]) <= 1e-5, nw


@pytest.mark.parametrize("precision,tol", [("double", 1e-12), ("mixed", 1e-6)])
def test_batch_matches_each_context(precision, tol):
    """Twelve systems of 32-200 waters through eight host threads give what
    each system's own context gives (PME grid atomics round in either
    order, hence the tolerance)."""
    systems = make_batch(count=12, lo=32, hi=200)
    print(f"[gpu] batch vs contexts {precision}", flush=True)
    with BatchEvaluator(systems, precision=precision, threads=8) as ev:
        a = ev.execute()
        b = ev.execute()
    for (f, pos, box), (ea, Fa), (eb, Fb) in zip(systems, a, b):
        ctx = MPIDContext(f, box, precision=precision)
        ctx.setPositions(pos)
        st = ctx.getState()
        e, F = st.getPotentialEnergy(), st.getForces()
        for ex, Fx in ((ea, Fa), (eb, Fb)):
            assert abs(ex - e) <= tol * abs(e)
            assert rel_rms(Fx, F) <= 10 * tol


def test_batch_errors_name_the_problem():
    f, pos, box = make_small_waterbox(32, seed=1)
    with pytest.raises(TypeError):
        BatchEvaluator([(None, pos, box)])
    with pytest.raises(ValueError):
        BatchEvaluator([(f, pos, box)], threads=0)
