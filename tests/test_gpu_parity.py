"""Parity of the gfx950 path (through the C ABI) with the FP64 oracle and with
the golden values of the reference's own tests.

Tolerances, stated per result:
  * DOUBLE precision path vs FP64 oracle: 1e-8 relative (same arithmetic,
    different summation order and solver).
  * MIXED precision path: force relative RMS <= 1e-5, energy relative <= 1e-6,
    induced-dipole relative RMS <= 1e-5 against the FP64 oracle (the targets of
    the workload definition), and the reference tests' own rule (1e-4) against
    their goldens.
"""
import numpy as np
import pytest

from mpidopenmmplugin_amd import MPIDForce, MPIDForceError
from mpidopenmmplugin_amd.kernel import MPIDContext, MPIDConvergenceError, HipCalcMPIDForceKernel
from mpidopenmmplugin_amd.waterbox import make_waterbox
from oracle import oracle
from tests.helpers import assert_equal_tol, assert_equal_vec, force_from_case, load_reference_tests

pytestmark = pytest.mark.gpu

CASES = load_reference_tests()["cases"]


def rel_rms(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return np.sqrt(((a - b) ** 2).sum() / (b ** 2).sum())


@pytest.mark.parametrize("precision", ["double", "mixed"])
@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_reference_goldens(case, precision, reference_tests):
    print(f"[gpu] golden {case['name']} {precision}", flush=True)
    f, pos, box = force_from_case(case, reference_tests["systems"])
    ctx = MPIDContext(f, box, precision=precision)
    ctx.setPositions(pos)
    st = ctx.getState()
    tol = case.get("energy_tol", 1e-4)
    assert_equal_tol(case["energy"], st.getPotentialEnergy(), tol)
    if "forces" in case:
        assert_equal_vec(case["forces"], st.getForces(), 1e-4)
    if precision == "double" and "forces" in case:
        ref = oracle.evaluate(f.flatten(), pos, box)
        assert abs(st.getPotentialEnergy() - ref["energy"]) <= 1e-8 * abs(ref["energy"])
        assert rel_rms(st.getForces(), ref["forces"]) <= 1e-8


@pytest.fixture(scope="module")
def water216():
    """The 648-atom workload; the oracle is converged to 1e-10 Debye so that
    solver noise does not enter the comparison."""
    f, pos, box = make_waterbox("water216")
    tight, _, _ = make_waterbox("water216", target_eps=1e-10)
    ref = oracle.evaluate(tight.flatten(), pos, box, details=True)
    return f, pos, box, ref


@pytest.mark.parametrize("precision,tol", [("double", 1e-10), ("mixed", 2e-6)])
def test_fixed_field(water216, precision, tol):
    print(f"[gpu] fixed field water216 {precision}", flush=True)
    f, pos, box, ref = water216
    k = HipCalcMPIDForceKernel(precision)
    k.initialize(f.getNumMultipoles(), f)
    E = k.computeFixedField(pos, box)
    assert rel_rms(E, ref["fixed_field"]) <= tol


@pytest.mark.parametrize("precision,tol", [("double", 1e-9), ("mixed", 2e-6)])
def test_induced_field_matvec(water216, precision, tol):
    print(f"[gpu] induced field water216 {precision}", flush=True)
    f, pos, box, ref = water216
    rng = np.random.default_rng(3)
    mu = rng.normal(size=pos.shape) * 2e-3
    want = oracle.induced_field(f.flatten(), pos, box, mu)
    k = HipCalcMPIDForceKernel(precision)
    k.initialize(f.getNumMultipoles(), f)
    got = k.computeInducedField(pos, box, mu)
    assert rel_rms(got, want) <= tol


@pytest.mark.parametrize("precision", ["double", "mixed"])
def test_full_evaluation_water216(water216, precision):
    """At the workload's own convergence target (1e-6 Debye): energy within
    1e-6, forces and induced dipoles within 1e-5 relative RMS of the FP64 oracle."""
    print(f"[gpu] full water216 {precision}", flush=True)
    f, pos, box, ref = water216
    ctx = MPIDContext(f, box, precision=precision)
    ctx.setPositions(pos)
    st = ctx.getState()
    e, F = st.getPotentialEnergy(), st.getForces()
    mu = ctx.getInducedDipoles()
    assert abs(e - ref["energy"]) <= 1e-6 * abs(ref["energy"])
    assert rel_rms(F, ref["forces"]) <= 1e-5
    assert rel_rms(mu, ref["induced"]) <= 1e-5
    lab = ctx.getLabFramePermanentDipoles()
    assert np.allclose(lab, ref["lab_dipoles"], rtol=0, atol=1e-12)
    tot = ctx.getTotalDipoles()
    assert np.allclose(tot, lab + mu, atol=1e-12)


def test_full_evaluation_water216_tight(water216):
    """FP64 path converged as tightly as the oracle: agreement to 1e-9."""
    f, pos, box, ref = water216
    g, _, _ = make_waterbox("water216", target_eps=1e-10)
    ctx = MPIDContext(g, box, precision="double")
    ctx.setPositions(pos)
    st = ctx.getState()
    assert abs(st.getPotentialEnergy() - ref["energy"]) <= 1e-9 * abs(ref["energy"])
    assert rel_rms(st.getForces(), ref["forces"]) <= 1e-8
    assert rel_rms(ctx.getInducedDipoles(), ref["induced"]) <= 1e-8


@pytest.mark.parametrize("precision,tol", [("double", 1e-12), ("mixed", 1e-6)])
def test_repeated_evaluations_agree(water216, precision, tol):
    """Per-atom sums are fixed point (order independent); the PME grid uses
    floating-point atomics, so repeated evaluations agree to rounding."""
    print(f"[gpu] repeat {precision}", flush=True)
    f, pos, box, ref = water216
    ctx = MPIDContext(f, box, precision=precision)
    ctx.setPositions(pos)
    a = ctx.getState()
    ctx.setPositions(pos + 0.0)
    b = ctx.getState()
    assert abs(a.getPotentialEnergy() - b.getPotentialEnergy()) <= tol * abs(a.getPotentialEnergy())
    assert rel_rms(a.getForces(), b.getForces()) <= 10 * tol


def test_forces_accumulate_into_buffer(water216):
    f, pos, box, ref = water216
    k = HipCalcMPIDForceKernel("double")
    k.initialize(f.getNumMultipoles(), f)
    base = np.ones((f.getNumMultipoles(), 3))
    e, out = k.execute(pos, box, forces=base.copy())
    e2, out2 = k.execute(pos, box)
    assert np.allclose(out, out2 + 1.0, atol=1e-9)


def test_nonconvergence_is_an_error(water216):
    f, pos, box, ref = water216
    f2, _, _ = make_waterbox("water216")
    f2.setMutualInducedMaxIterations(1)
    f2.setMutualInducedTargetEpsilon(1e-12)
    ctx = MPIDContext(f2, box, precision="double")
    ctx.setPositions(pos)
    with pytest.raises(MPIDConvergenceError, match="did not converge"):
        ctx.getState()


@pytest.mark.parametrize("precision", ["double", "mixed"])
def test_getters_after_failed_evaluation(water216, precision):
    """A failed evaluation must not leave a getter serving the last good
    one's cache: after execute(A) succeeds and execute(B) fails to converge
    (it has rewritten the sorted positions, permutation and dipoles), the
    induced dipoles at A equal those of a fresh context."""
    print(f"[gpu] getters after a failed evaluation {precision}", flush=True)
    f, pos, box, _ = water216
    rng = np.random.default_rng(11)
    # two configurations whose solves take different iteration counts
    configs = [pos] + [pos + rng.normal(scale=s, size=pos.shape) for s in (0.004, 0.008, 0.012, 0.016, 0.02)]
    probe = MPIDContext(f, box, precision=precision)
    its = []
    for c in configs:
        probe.setPositions(c)
        probe.getState()
        its.append(probe.kernel.stats()["iterations"])
    a, b = int(np.argmin(its)), int(np.argmax(its))
    assert its[b] > its[a], its
    fresh = MPIDContext(f, box, precision=precision)
    fresh.setPositions(configs[a])
    fresh.getState()
    want = fresh.getInducedDipoles()
    g, _, _ = make_waterbox("water216")
    g.setMutualInducedMaxIterations(its[a])
    ctx = MPIDContext(g, box, precision=precision)
    ctx.setPositions(configs[a])
    ctx.getState()
    ctx.setPositions(configs[b])
    with pytest.raises(MPIDConvergenceError):
        ctx.getState()
    ctx.setPositions(configs[a])
    got = ctx.getInducedDipoles()
    assert rel_rms(got, want) <= (1e-12 if precision == "double" else 1e-6)


def test_invalid_parameters_raise():
    f, pos, box = make_waterbox("water216")
    f.setCutoffDistance(1.0)  # more than half of the 1.862 nm box
    ctx = MPIDContext(f, box, precision="double")
    ctx.setPositions(pos)
    with pytest.raises(MPIDForceError, match="twice the nonbonded cutoff"):
        ctx.getState()
    g, pos, box = make_waterbox("water216")
    q = list(g.getMultipoleParameters(0))
    q[2] = [1e-3, 0, 0, 0, 0, 0]  # non-traceless quadrupole
    g.setMultipoleParameters(0, *q)
    with pytest.raises(MPIDForceError, match="nonzero trace"):
        MPIDContext(g, box, precision="double")


def test_update_parameters_in_context(water216):
    """copyParametersToContext (without the reference platform's octopole bug,
    platforms/reference/src/MPIDReferenceKernels.cpp:371-376)."""
    f, pos, box, ref = water216
    g, _, _ = make_waterbox("water216", target_eps=1e-10)
    ctx = MPIDContext(g, box, precision="double")
    ctx.setPositions(pos)
    e0 = ctx.getState().getPotentialEnergy()
    q = list(g.getMultipoleParameters(0))
    q[0] = -0.7
    g.setMultipoleParameters(0, *q)
    ctx.updateParametersInContext()
    e1 = ctx.getState().getPotentialEnergy()
    want = oracle.evaluate(g.flatten(), pos, box)["energy"]
    assert abs(e1 - want) <= 1e-9 * abs(want)
    assert abs(e0 - ref["energy"]) <= 1e-9 * abs(ref["energy"])


@pytest.fixture(scope="module")
def water4096():
    import os
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "water4096_oracle.npz")
    ref = np.load(path)
    f, pos, box = make_waterbox("water4096")
    assert abs(pos.sum() - float(ref["positions_checksum"])) < 1e-9, "generator drifted from the fixture"
    return f, pos, box, ref


@pytest.mark.parametrize("precision", ["double", "mixed"])
def test_water4096_octopoles_anisotropic(water4096, precision):
    """12,288 atoms, octopoles and anisotropic polarisability, 64^3 grid,
    against the committed FP64 oracle fixture (converged to 1e-10 Debye)."""
    print(f"[gpu] water4096 {precision}", flush=True)
    f, pos, box, ref = water4096
    ctx = MPIDContext(f, box, precision=precision)
    ctx.setPositions(pos)
    st = ctx.getState()
    e_ref = float(ref["energy"])
    assert abs(st.getPotentialEnergy() - e_ref) <= 1e-6 * abs(e_ref)
    assert rel_rms(st.getForces(), ref["forces"]) <= 1e-5
    assert rel_rms(ctx.getInducedDipoles(), ref["induced"]) <= 1e-5


@pytest.fixture(scope="module")
def water32768():
    import os
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "water32768_oracle.npz")
    ref = np.load(path)
    f, pos, box = make_waterbox("water32768")
    assert abs(pos.sum() - float(ref["positions_checksum"])) < 1e-8, "generator drifted from the fixture"
    return f, pos, box, ref


@pytest.mark.parametrize("precision", ["double", "mixed"])
def test_water32768(water32768, precision):
    """The 98,304-atom benchmark box (128^3 grid, 0.9 nm cutoff, octopoles,
    anisotropic polarisability) against the committed FP64 oracle fixture
    (converged to 1e-10 Debye).  The FP64 path, converged to 1e-10 Debye
    too, agrees to 1e-8; the mixed path at the workload's 1e-6 Debye meets
    energy 1e-6 and force and induced-dipole RMS 1e-5.  This is the only
    size that runs the 128^3 FFT passes (the fused FP32 plane pass and the
    128-point z transforms in both precisions)."""
    print(f"[gpu] water32768 {precision}", flush=True)
    f, pos, box, ref = water32768
    if precision == "double":
        f, _, _ = make_waterbox("water32768", target_eps=1e-10)
        tol_e, tol_f = 1e-8, 1e-8
    else:
        tol_e, tol_f = 1e-6, 1e-5
    ctx = MPIDContext(f, box, precision=precision)
    ctx.setPositions(pos)
    st = ctx.getState()
    e_ref = float(ref["energy"])
    e_err = abs(st.getPotentialEnergy() - e_ref) / abs(e_ref)
    f_err = rel_rms(st.getForces(), ref["forces"])
    mu_err = rel_rms(ctx.getInducedDipoles(), ref["induced"])
    print(f"[gpu] water32768 {precision}: energy {e_err:.2e}  force RMS {f_err:.2e}  dipole RMS {mu_err:.2e}",
          flush=True)
    assert e_err <= tol_e
    assert f_err <= tol_f
    assert mu_err <= tol_f


@pytest.mark.parametrize("conservative", [False, True], ids=["reference_torques", "conservative_torques"])
@pytest.mark.parametrize("precision", ["double", "mixed"])
def test_water216x_octopoles_anisotropic(precision, conservative):
    """Octopoles and anisotropic polarisability (the parameter set of the
    benchmark box) on 648 atoms, in both torque conventions."""
    print(f"[gpu] water216x {precision} conservative={conservative}", flush=True)
    tight, pos, box = make_waterbox("water216x", target_eps=1e-10)
    ref = oracle.evaluate(tight.flatten(), pos, box, details=True, conservative_torques=conservative)
    if precision == "double":
        ctx = MPIDContext(tight, box, precision="double", conservative_torques=conservative)
        e_tol, f_tol = 1e-9, 1e-8
    else:
        f, _, _ = make_waterbox("water216x")
        ctx = MPIDContext(f, box, precision="mixed", conservative_torques=conservative)
        e_tol, f_tol = 1e-6, 1e-5
    ctx.setPositions(pos)
    st = ctx.getState()
    assert abs(st.getPotentialEnergy() - ref["energy"]) <= e_tol * abs(ref["energy"])
    assert rel_rms(st.getForces(), ref["forces"]) <= f_tol
    assert rel_rms(ctx.getInducedDipoles(), ref["induced"]) <= f_tol


@pytest.mark.parametrize("precision", ["double", "mixed"])
def test_graph_replays_match_eager_iterations(water4096, precision):
    """The solver's iterations run as replays of a captured graph unless
    per-kernel profiling is on, when they are launched one by one.  Both give
    the same result: an eager evaluation, then two graph evaluations in a row
    (the second replays an executable graph updated in place), each compared
    with the eager one.  The tolerance covers the order of the FP64 atomic
    sums in the spreading kernels, which is not fixed from run to run."""
    print(f"[gpu] graph vs eager water4096 {precision}", flush=True)
    f, pos, box, _ = water4096
    k = HipCalcMPIDForceKernel(precision)
    k.initialize(len(pos), f)
    k.setProfiling(True)
    e0, f0 = k.execute(pos, box)
    it0 = k.stats()["iterations"]
    k.setProfiling(False)
    tol_e, tol_f = (1e-12, 1e-10) if precision == "double" else (1e-9, 1e-6)
    for rep in range(2):
        e1, f1 = k.execute(pos, box)
        assert k.stats()["iterations"] == it0, rep
        assert abs(e1 - e0) <= tol_e * abs(e0), rep
        assert rel_rms(f1, f0) <= tol_f, rep


@pytest.mark.parametrize("precision", ["double", "mixed"])
def test_iteration_limit_boundary(water216, precision):
    """Each solver iteration is issued before the host has read the previous
    iteration's convergence check.  With the iteration limit set to exactly
    the number of iterations the solve takes, it converges with that count
    and the same result; one fewer is an error.  The contexts spread with
    the library's tiling (windows summed in FP64 in LDS), whose reruns agree
    to 1e-9 in mixed precision; without tiles (the autotuned choice for this
    24^3 grid) the FP32 grid atomics vary the mixed energy by a few 1e-9."""
    print(f"[gpu] iteration limit {precision}", flush=True)
    f, pos, box, _ = water216
    k = HipCalcMPIDForceKernel(precision, spread_tiling="library")
    k.initialize(len(pos), f)
    e0, f0 = k.execute(pos, box)
    it0 = k.stats()["iterations"]
    assert it0 >= 2
    tol_e, tol_f = (1e-12, 1e-10) if precision == "double" else (1e-9, 1e-6)
    g, _, _ = make_waterbox("water216")
    g.setMutualInducedMaxIterations(it0)
    k1 = HipCalcMPIDForceKernel(precision, spread_tiling="library")
    k1.initialize(len(pos), g)
    e1, f1 = k1.execute(pos, box)
    assert k1.stats()["iterations"] == it0
    assert abs(e1 - e0) <= tol_e * abs(e0)
    assert rel_rms(f1, f0) <= tol_f
    h, _, _ = make_waterbox("water216")
    h.setMutualInducedMaxIterations(it0 - 1)
    k2 = HipCalcMPIDForceKernel(precision, spread_tiling="library")
    k2.initialize(len(pos), h)
    with pytest.raises(MPIDConvergenceError, match="did not converge"):
        k2.execute(pos, box)


@pytest.mark.parametrize("precision", ["double", "mixed"])
def test_list_overflow_retry(water4096, precision):
    """After its first evaluation a context does not wait for the
    neighbour-list build; the overflow check is read at the solver's first
    synchronisation, with a solver iteration already issued.  Compressing the
    box by 0.85 per edge (1.63 times the density) overflows the lists sized
    for the first density: the evaluation is repeated with larger lists and
    must give what a fresh context gives for the compressed box."""
    print(f"[gpu] list overflow retry {precision}", flush=True)
    f, pos, box, _ = water4096
    s = 0.85
    k = HipCalcMPIDForceKernel(precision)
    k.initialize(len(pos), f)
    k.execute(pos, box)
    e1, f1 = k.execute(pos * s, np.asarray(box) * s)
    fresh = HipCalcMPIDForceKernel(precision)
    fresh.initialize(len(pos), f)
    e2, f2 = fresh.execute(pos * s, np.asarray(box) * s)
    tol_e, tol_f = (1e-12, 1e-10) if precision == "double" else (1e-9, 1e-6)
    assert k.stats()["iterations"] == fresh.stats()["iterations"]
    assert abs(e1 - e2) <= tol_e * abs(e2)
    assert rel_rms(f1, f2) <= tol_f


@pytest.mark.parametrize("single_stream", [False, True], ids=["three_streams", "single_stream"])
def test_mixed_refinement_tight_target(water216, single_stream):
    """A target beyond FP32 arithmetic (1e-8 Debye) makes the mixed path
    refine: the FP64 residual of the FP32 solution is above the target, so a
    correction is solved and the final pass, issued beside the check, is
    issued again.  The result must then be as close to the FP64 oracle
    (converged to 1e-10) as the target allows, and the first final pass must
    leave no trace in the outputs (forces accumulate into a buffer of ones).
    With single_stream the check runs after the final pass on the same
    stream instead of beside it."""
    print(f"[gpu] mixed refinement water216 single_stream={single_stream}", flush=True)
    f, pos, box, ref = water216
    g, _, _ = make_waterbox("water216", target_eps=1e-8)
    k = HipCalcMPIDForceKernel("mixed", single_stream=single_stream)
    k.initialize(len(pos), g)
    base = np.ones((len(pos), 3))
    e, F = k.execute(pos, box, forces=base.copy())
    st = k.stats()
    assert st["converged"] and st["refinements"] >= 1, st
    assert st["epsilon"] < 1e-8
    assert abs(e - ref["energy"]) <= 1e-7 * abs(ref["energy"])
    assert rel_rms(F - 1.0, ref["forces"]) <= 1e-6
    assert rel_rms(k.getInducedDipoles(pos, box), ref["induced"]) <= 1e-6


@pytest.mark.parametrize("precision", ["double", "mixed"])
def test_extrapolated_water216x(precision):
    """Extrapolated (OPT) polarization with the default coefficients on 648
    atoms with octopoles and anisotropic polarisability (so the response
    torques enter), against the FP64 oracle: 1e-9 energy and 1e-8 forces in
    FP64, 1e-6 and 1e-5 in mixed precision."""
    print(f"[gpu] extrapolated water216x {precision}", flush=True)
    f, pos, box = make_waterbox("water216x", polarization=MPIDForce.Extrapolated)
    ref = oracle.evaluate(f.flatten(), pos, box, details=True)
    ctx = MPIDContext(f, box, precision=precision)
    ctx.setPositions(pos)
    st = ctx.getState()
    e_tol, f_tol = (1e-9, 1e-8) if precision == "double" else (1e-6, 1e-5)
    assert abs(st.getPotentialEnergy() - ref["energy"]) <= e_tol * abs(ref["energy"])
    assert rel_rms(st.getForces(), ref["forces"]) <= f_tol
    assert rel_rms(ctx.getInducedDipoles(), ref["induced"]) <= f_tol


@pytest.mark.parametrize("polarization", [MPIDForce.Mutual, MPIDForce.Extrapolated], ids=["mutual", "extrapolated"])
@pytest.mark.parametrize("precision", ["double", "mixed"])
def test_nocutoff_water216x(precision, polarization):
    """NoCutoff on 648 atoms with octopoles and anisotropic polarisability:
    every pair interacts, with no periodic images, no reciprocal and no self
    terms.  The FP64 oracle is the reference here (its NoCutoff mode is pinned
    by the reference's six NoCutoff goldens).  The box passed in is ignored."""
    print(f"[gpu] nocutoff water216x {precision} pol={polarization}", flush=True)
    f, pos, box = make_waterbox("water216x", polarization=polarization, target_eps=1e-10)
    f.setNonbondedMethod(MPIDForce.NoCutoff)
    ref = oracle.evaluate(f.flatten(), pos, box, details=True)
    g, _, _ = make_waterbox("water216x", polarization=polarization)
    g.setNonbondedMethod(MPIDForce.NoCutoff)
    k = HipCalcMPIDForceKernel(precision)
    k.initialize(len(pos), g if precision == "mixed" else f)
    e, F = k.execute(pos, None)
    e_tol, f_tol = (1e-9, 1e-8) if precision == "double" else (1e-6, 1e-5)
    assert abs(e - ref["energy"]) <= e_tol * abs(ref["energy"])
    assert rel_rms(F, ref["forces"]) <= f_tol
    assert rel_rms(k.getInducedDipoles(pos, None), ref["induced"]) <= f_tol


@pytest.mark.parametrize("method", ["pme", "nocutoff"])
@pytest.mark.parametrize("precision", ["double", "mixed"])
def test_potential_and_moments_water216x(precision, method):
    """getElectrostaticPotential at 64 points and getSystemMultipoleMoments
    after an evaluation, against the oracle (whose forms are pinned in
    tests/test_oracle_getters.py)."""
    print(f"[gpu] potential and moments water216x {precision} {method}", flush=True)
    tight, pos, box = make_waterbox("water216x", target_eps=1e-10)
    f, _, _ = make_waterbox("water216x")
    if method == "nocutoff":
        tight.setNonbondedMethod(MPIDForce.NoCutoff)
        f.setNonbondedMethod(MPIDForce.NoCutoff)
    rng = np.random.default_rng(11)
    pts = rng.uniform(0.0, float(np.asarray(box).reshape(-1)[0]), (64, 3))
    masses = np.tile([15.999, 1.008, 1.008], len(pos) // 3)
    ref_phi = oracle.electrostatic_potential(tight.flatten(), pos, box, pts)
    ref_mom = oracle.system_multipole_moments(tight.flatten(), pos, box, masses)
    k = HipCalcMPIDForceKernel(precision)
    k.initialize(len(pos), tight if precision == "double" else f)
    phi = k.getElectrostaticPotential(pts, pos, box)
    mom = k.getSystemMultipoleMoments(pos, box, masses)
    tol = 1e-9 if precision == "double" else 1e-5
    assert rel_rms(phi, ref_phi) <= tol
    assert abs(mom[0] - ref_mom[0]) <= 1e-9
    assert rel_rms(mom[1:], ref_mom[1:]) <= tol
