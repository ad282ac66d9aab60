"""Parity at the edges of the drop-in boundary, through the C ABI, against
the FP64 oracle: molecules with many scaled covalent partners per atom
(MPIDReferenceForce.cpp:190-249 keeps them in per-atom maps without a
limit), and reduced triclinic boxes (MPIDReferenceForce.cpp:2615-2633,
2671-2676, 3077-3169).

Tolerances: FP64 path 1e-8 relative in energy and force RMS; mixed path
energy 1e-6, force and induced-dipole RMS 1e-5 (the workload targets)."""
import numpy as np
import pytest

from mpidopenmmplugin_amd import MPIDForce
from mpidopenmmplugin_amd.kernel import MPIDContext
from oracle import oracle
from mpidopenmmplugin_amd.waterbox import make_waterbox
from tests.helpers import make_branched_cluster, sheared_waters, triclinic_supercell

pytestmark = pytest.mark.gpu


def rel_rms(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return np.sqrt(((a - b) ** 2).sum() / (b ** 2).sum())


def _compare(f, pos, box, precision, ref):
    ctx = MPIDContext(f, box, precision=precision)
    ctx.setPositions(pos)
    st = ctx.getState()
    e_tol, f_tol = (1e-8, 1e-8) if precision == "double" else (1e-6, 1e-5)
    assert abs(st.getPotentialEnergy() - ref["energy"]) <= e_tol * abs(ref["energy"])
    assert rel_rms(st.getForces(), ref["forces"]) <= f_tol
    assert rel_rms(ctx.getInducedDipoles(), ref["induced"]) <= f_tol


@pytest.mark.parametrize("scale14", [0.5, 1.0])
@pytest.mark.parametrize("method", ["nocutoff", "pme"])
@pytest.mark.parametrize("precision", ["double", "mixed"])
def test_many_covalent_partners(precision, method, scale14):
    """The centre of a 61-atom tree has 60 scaled partners (1-2, 1-3 and
    1-4), far beyond 16 slots; at scale14 = 1 the 1-4 pairs are ordinary
    pairs.  Octopoles and anisotropic polarisabilities on every site."""
    print(f"[gpu] many partners {precision} {method} scale14={scale14}", flush=True)
    f, pos, npart = make_branched_cluster(scale14=scale14)
    assert npart > 16
    f.setPolarizationType(MPIDForce.Mutual)
    box = None
    if method == "pme":
        f.setNonbondedMethod(MPIDForce.PME)
        f.setCutoffDistance(0.9)
        f.setPMEParameters(3.2, 32, 32, 32)
        box = np.diag([2.4, 2.4, 2.4])
    else:
        f.setNonbondedMethod(MPIDForce.NoCutoff)
    f.setMutualInducedTargetEpsilon(1e-10)
    ref = oracle.evaluate(f.flatten(), pos, box, details=True)
    if precision == "mixed":
        f.setMutualInducedTargetEpsilon(1e-6)
    _compare(f, pos, box, precision, ref)


# generic tilts (fractions of the edge): b = (0.25 L, L, 0), c = (-0.30 L, 0.20 L, L).
# The cell's smallest height is then 0.92 L (1.71 nm for the 216-water box),
# so the 0.8 nm cutoff stays below half of every height and no pair reaches
# two images of the same atom.
TILTS = (0.25, -0.30, 0.20)


def _sheared_box(name):
    f, pos, box = make_waterbox(name)
    L = box[0]
    vecs = np.array([[L, 0.0, 0.0], [TILTS[0] * L, L, 0.0], [TILTS[1] * L, TILTS[2] * L, L]])
    return f, sheared_waters(pos, L, vecs), vecs


_REF = {}


def _oracle_ref(key, f, pos, box):
    if key not in _REF:
        flat = f.flatten()
        flat["target_eps"] = 1e-10
        _REF[key] = oracle.evaluate(flat, pos, box, details=True)
    return _REF[key]


@pytest.mark.parametrize("precision", ["double", "mixed"])
def test_triclinic_sheared_waters(precision):
    """216 waters (octopoles, anisotropic polarisabilities) moved molecule by
    molecule into a reduced triclinic cell with three generic tilts: energy,
    forces and induced dipoles, and the potential at points inside and
    outside the cell."""
    print(f"[gpu] triclinic sheared water216x {precision}", flush=True)
    f, pos, vecs = _sheared_box("water216x")
    ref = _oracle_ref("sheared216", f, pos, vecs)
    f.setMutualInducedTargetEpsilon(1e-10 if precision == "double" else 1e-6)
    _compare(f, pos, vecs, precision, ref)
    rng = np.random.default_rng(3)
    pts = rng.uniform(-0.2, 1.2, (200, 3)) @ vecs
    want = oracle.electrostatic_potential(dict(f.flatten(), target_eps=1e-10), pos, vecs, pts)
    ctx = MPIDContext(f, vecs, precision=precision)
    ctx.setPositions(pos)
    got = ctx.getElectrostaticPotential(pts)
    tol = 1e-8 if precision == "double" else 1e-5
    assert rel_rms(got, want) <= tol


@pytest.mark.parametrize("precision", ["double", "mixed"])
def test_triclinic_supercell(precision):
    """Four copies of the 216-water cell in the triclinic cell (2L,0,0),
    (L,2L,0), (L,L,L), every off-diagonal at its largest allowed value
    (2,592 atoms, 96 x 96 x 48 grid), with the copies also moved by whole
    lattice vectors (positions far outside the cell)."""
    print(f"[gpu] triclinic supercell {precision}", flush=True)
    f, pos, box = make_waterbox("water216x")
    L = box[0]
    alpha = f.getPMEParameters()[0]
    f4, p4, vecs = triclinic_supercell(f, pos, L, alpha, 48)
    n = len(pos)
    ref = _oracle_ref("supercell", f4, p4, vecs)
    f4.setMutualInducedTargetEpsilon(1e-10 if precision == "double" else 1e-6)
    _compare(f4, p4, vecs, precision, ref)
    moved = p4.copy()
    moved[n:2 * n] += vecs[1]
    moved[2 * n:3 * n] -= 2.0 * vecs[2]
    moved[3 * n:] += vecs[0] + vecs[1] - vecs[2]
    _compare(f4, moved, vecs, precision, ref)


@pytest.mark.parametrize("precision", ["double", "mixed"])
def test_triclinic_water4096(precision):
    """The 12,288-atom box sheared with the same tilts: the tiled spreading
    and the multi-tile neighbour lists on a triclinic cell."""
    print(f"[gpu] triclinic sheared water4096 {precision}", flush=True)
    f, pos, vecs = _sheared_box("water4096")
    ref = _oracle_ref("sheared4096", f, pos, vecs)
    f.setMutualInducedTargetEpsilon(1e-10 if precision == "double" else 1e-6)
    _compare(f, pos, vecs, precision, ref)


def test_box_not_in_reduced_form_is_rejected():
    """OpenMM keeps box vectors in reduced form; others are an error."""
    f, pos, box = make_waterbox("water216")
    L = box[0]
    for bad in ([[L, 0, 0], [0.6 * L, L, 0], [0, 0, L]],    # |bx| > ax / 2
                [[L, 0, 0], [0, L, 0], [0, 0.6 * L, L]],    # |cy| > by / 2
                [[L, 0.1, 0], [0, L, 0], [0, 0, L]]):       # a not along x
        ctx = MPIDContext(f, np.array(bad), precision="double")
        with pytest.raises(Exception, match="reduced form"):
            ctx.setPositions(pos)
            ctx.getState()


@pytest.mark.parametrize("grid,tilt", [(15, 0.0), (15, 1.0), (16, 1.0)], ids=["odd", "odd-triclinic", "even-triclinic"])
@pytest.mark.parametrize("precision", ["double", "mixed"])
def test_influence_function_asymmetric_planes(precision, grid, tilt):
    """The planes where the reference's influence function is not symmetric
    in k -> -k (performMPIDReciprocalConvolution multiplies the full complex
    spectrum and keeps the real part): k = (n - 1) / 2 of an odd grid, whose
    optimal-zeta factor folds asymmetrically (MPIDReferenceForce.cpp:2780-2806),
    and the Nyquist planes of a triclinic cell, where the fold keeps +n/2.  A
    large Ewald alpha on a coarse grid (32 waters, 0.35 nm cutoff, alpha 7.5)
    gives those planes weight: multiplying by G(k) instead of its symmetric
    part moves the energy by 5e-8 to 1e-3 relative here."""
    from mpidopenmmplugin_amd.waterbox import ewald_alpha, make_small_waterbox
    print(f"[gpu] influence planes grid {grid} tilt {tilt} {precision}", flush=True)
    f, pos, box = make_small_waterbox(32, seed=7)
    f.setCutoffDistance(0.35)
    f.setPMEParameters(ewald_alpha(0.35), grid, grid, grid)
    L = box[0]
    vecs = np.array([[L, 0.0, 0.0], [0.31 * L * tilt, L, 0.0], [-0.42 * L * tilt, 0.27 * L * tilt, L]])
    if tilt:
        pos = sheared_waters(pos, L, vecs)
    ref = _oracle_ref(("planes", grid, tilt), f, pos, vecs)
    f.setMutualInducedTargetEpsilon(1e-10 if precision == "double" else 1e-6)
    _compare(f, pos, vecs, precision, ref)
