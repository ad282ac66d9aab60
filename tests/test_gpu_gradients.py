"""Energy gradients (mpid_energy_gradients, through the C ABI) against central
finite differences of the FP64 oracle's energy, re-converged at every point.

The derivative of the energy in a parameter is fixed by the energy itself, so
differences of the oracle's energy pin every gradient the library returns:
positions, charges, molecular-frame dipoles, quadrupoles and octopoles (each
listed component), polarisabilities, Thole factors and the default Thole
width.  The differences take the oracle's variational energy (the energy
plus -1/2 eps mu.r, r the solver's residual), whose error is second order
in the solver's: the plain energy moves by ~1e-7 kJ/mol between solves,
which would swamp the small Thole derivatives.  The energy is quadratic in
the multipole parameters, so their two-point differences carry no
truncation error; positions, polarisabilities and Thole factors take the
four-point stencil (error O(h^4)).

Tolerances, per parameter class, relative to the largest gradient of the
class: DOUBLE (converged to 1e-10 Debye) 1e-6; MIXED (1e-6 Debye, FP32 pair
and reciprocal arithmetic) 1e-4, with the per-type sums of the 12,288-atom
box at 1e-5 of each gradient (the workload's bar).
"""
import os

import numpy as np
import pytest

from mpidopenmmplugin_amd import MPIDForce
from mpidopenmmplugin_amd.kernel import HipCalcMPIDForceKernel, MPIDUnsupportedError
from mpidopenmmplugin_amd.waterbox import make_small_waterbox, make_waterbox
from oracle import oracle
from tests.helpers import frames_cluster

pytestmark = pytest.mark.gpu

# absolute difference steps per class; the classes in which the energy is not
# quadratic take the four-point stencil
STEPS = dict(charges=1e-5, dipoles=1e-6, quadrupoles=1e-7, octopoles=1e-8, polarity=1e-5, tholes=1e-3,
             positions=1e-4)
FOUR_POINT = ("positions", "polarity", "tholes")
WIDTH = dict(charges=1, dipoles=3, quadrupoles=6, octopoles=10, polarity=3, tholes=1)


def _energy(flat, pos, box):
    return oracle.evaluate(flat, pos, box)["energy_variational"]


def _diff(f, h, four):
    """d/dx of f at 0 from f(+-h) (and f(+-2h) with four)."""
    if four:
        return (8.0 * (f(h) - f(-h)) - (f(2 * h) - f(-2 * h))) / (12.0 * h)
    return (f(h) - f(-h)) / (2.0 * h)


def _kernel(force, box, precision):
    k = HipCalcMPIDForceKernel(precision)
    k.initialize(force.getNumMultipoles(), force, default_box=box)
    return k


def fd_gradients(flat, pos, box, sites, classes, target_eps=1e-10):
    """Central differences of the oracle's energy in the listed per-site
    parameters: {class: [len(sites), width]}."""
    flat = dict(flat)
    flat["target_eps"] = target_eps
    out = {}
    for cls in classes:
        w = 3 if cls == "positions" else WIDTH[cls]
        g = np.zeros((len(sites), w))
        for a, site in enumerate(sites):
            for c in range(w):
                def shifted(x, site=site, c=c):
                    f2 = dict(flat)
                    p2 = pos
                    if cls == "positions":
                        p2 = pos.copy()
                        p2[site, c] += x
                    else:
                        arr = np.array(flat[cls], dtype=np.float64).reshape(-1, w)
                        arr[site, c] += x
                        f2[cls] = arr.reshape(-1)
                    return _energy(f2, p2, box)
                g[a, c] = _diff(shifted, STEPS[cls], cls in FOUR_POINT)
        out[cls] = g
    return out


def fd_default_thole(flat, pos, box, h=2e-2, target_eps=1e-10):
    """Four-point differences in the default Thole width.  Its derivative is
    small (~3e-3 kJ/mol per unit width on the 96-atom box) against the
    rounding of the energy: at h = 1e-4 to 1e-3 the estimates still move by
    1e-5 relative from step to step, from h = 3e-3 to 3e-2 they agree to
    1e-6 (and the O(h^4) error is below that)."""
    def shifted(x):
        f2 = dict(flat)
        f2["target_eps"] = target_eps
        f2["default_thole"] = flat["default_thole"] + x
        return _energy(f2, pos, box)
    return _diff(shifted, h, True)


# the frames cluster's sites (ThreeFold, ZBisect, ZThenX+y, Bisector, ZOnly,
# ZThenX (mirror), NoAxisType) and the classes compared at them; their
# differences are the committed fixture frames_gradients_fd.npz
# (tests/golden/make_frames_fd_fixture.py: about 400 oracle evaluations)
FRAME_SITES = [0, 1, 2, 3, 7, 10, 15]
FRAME_CLASSES = ["positions", "charges", "dipoles", "quadrupoles", "octopoles", "polarity", "tholes"]
FRAME_INPUTS = ["charges", "dipoles", "quadrupoles", "octopoles", "polarity", "tholes", "axis_types", "atom_z",
                "atom_x", "atom_y", "cov_offsets", "cov_atoms", "alpha", "cutoff", "grid", "default_thole", "scale14",
                "polarization", "nonbonded_method"]


def frame_box(box):
    return np.zeros(9) if box is None else np.asarray(box, dtype=np.float64).reshape(-1)
FRAMES_FD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "frames_gradients_fd.npz")


def frames_fd(case, flat, pos, box):
    """The oracle's differences for the frames cluster from the fixture,
    after checking that they were taken at these positions, box and
    parameters."""
    with np.load(FRAMES_FD) as z:
        np.testing.assert_array_equal(z[f"{case}_in_positions"], np.asarray(pos, dtype=np.float64))
        np.testing.assert_array_equal(z[f"{case}_in_box"], frame_box(box))
        for name in FRAME_INPUTS:
            np.testing.assert_array_equal(z[f"{case}_in_{name}"], np.asarray(flat[name], dtype=np.float64),
                                          err_msg=name)
        return {cls: z[f"{case}_{cls}"] for cls in FRAME_CLASSES}


_FD_CACHE = {}


def cached(key, make):
    """The finite differences of a case, computed once for both precisions."""
    if key not in _FD_CACHE:
        _FD_CACHE[key] = make()
    return _FD_CACHE[key]


def compare(got, fd, sites, tol, what):
    for cls, ref in fd.items():
        arr = np.asarray(got[cls])
        g = arr.reshape(arr.shape[0], -1)[sites]
        scale = np.abs(ref).max()
        err = np.abs(g - ref).max() / scale
        print(f"  {what} {cls:12s} max |g| {scale:.3e}  max rel err {err:.2e}", flush=True)
        assert scale > 0.0, cls
        assert err <= tol, (cls, err, g, ref)


@pytest.mark.parametrize("precision", ["double", "mixed"])
def test_small_water_box_every_parameter(precision):
    """96 atoms (32 waters, 15^3 grid: the untiled spreading path), octopoles
    and anisotropic polarisabilities; two oxygens and one hydrogen, every
    parameter, with the bare hydrogens' moments through all_sites."""
    print(f"[gpu] gradients small water box {precision}", flush=True)
    force, pos, box = make_small_waterbox(32, seed=7)
    force.setMutualInducedTargetEpsilon(1e-10 if precision == "double" else 1e-6)
    flat = force.flatten()
    sites = [0, 1, 15]

    def make():
        fd = fd_gradients(flat, pos, box, sites, ["positions", "charges", "dipoles", "quadrupoles", "octopoles"])
        fd.update(fd_gradients(flat, pos, box, [0, 15], ["polarity"]))
        fd["default_thole"] = fd_default_thole(flat, pos, box)
        fd["energy"] = oracle.evaluate(dict(flat, target_eps=1e-10), pos, box)["energy_variational"]
        return fd
    fd = dict(cached("small", make))
    gdef = fd.pop("default_thole")
    ref = fd.pop("energy")
    k = _kernel(force, box, precision)
    got = k.energyGradients(pos, box, all_sites=True)
    tol = 1e-6 if precision == "double" else 1e-4
    compare(got, {c: v for c, v in fd.items() if c != "polarity"}, sites, tol, precision)
    compare(got, {"polarity": fd["polarity"]}, [0, 15], tol, precision)
    print(f"  default Thole width: {got['default_thole']:.9e} fd {gdef:.9e}", flush=True)
    assert abs(got["default_thole"] - gdef) <= tol * abs(gdef)
    # the energy of the refined dipoles
    assert abs(got["energy"] - ref) <= (1e-9 if precision == "double" else 1e-6) * abs(ref)
    # without all_sites: the bare hydrogens keep their charge gradient, the rest is zero
    fast = k.energyGradients(pos, box)
    assert np.allclose(fast["charges"], got["charges"], rtol=1e-9 if precision == "double" else 1e-5, atol=1e-9)
    assert np.all(fast["dipoles"][1::3] == 0.0) and np.all(fast["octopoles"][2::3] == 0.0)
    # two solves of their own: equal to the solver's tolerance
    rt = 1e-6 if precision == "double" else 1e-4
    np.testing.assert_allclose(fast["dipoles"][0::3], got["dipoles"][0::3], rtol=rt, atol=1e-9)
    np.testing.assert_allclose(fast["polarity"][0::3], got["polarity"][0::3], rtol=rt, atol=1e-12)
    assert fast["adjoint_iterations"] >= 0


@pytest.mark.parametrize("periodic", [False, True], ids=["nocutoff", "pme"])
@pytest.mark.parametrize("precision", ["double", "mixed"])
def test_every_frame_type(precision, periodic):
    """The rotation back to the molecular frame for ThreeFold, ZBisect,
    ZThenX (with a y atom, both handednesses), Bisector, ZOnly and NoAxisType
    sites, and their position gradients (ThreeFold torques by the exact
    chain rule); every site polarisable, so the intramolecular pairs
    (polarization scale 0) carry the per-site Thole factors."""
    print(f"[gpu] gradients frames cluster {precision} periodic={periodic}", flush=True)
    force, pos, box = frames_cluster(periodic=periodic, target_eps=1e-10)
    flat = force.flatten()
    if precision == "mixed":
        force.setMutualInducedTargetEpsilon(1e-8)
    fd = frames_fd("pme" if periodic else "nocutoff", flat, pos, box)
    got = _kernel(force, box, precision).energyGradients(pos, box)
    compare(got, fd, FRAME_SITES, 1e-6 if precision == "double" else 1e-4, precision)


@pytest.mark.parametrize("precision", ["double", "mixed"])
def test_direct_polarization(precision):
    print(f"[gpu] gradients direct polarization {precision}", flush=True)
    force, pos, box = make_small_waterbox(32, seed=9, polarization=MPIDForce.Direct)
    sites = [3, 4]

    def make():
        fd = fd_gradients(force.flatten(), pos, box, sites, ["positions", "charges", "dipoles", "octopoles"])
        fd.update(fd_gradients(force.flatten(), pos, box, [3], ["polarity"]))
        return fd
    fd = cached("direct", make)
    got = _kernel(force, box, precision).energyGradients(pos, box, all_sites=True)
    tol = 1e-6 if precision == "double" else 1e-4
    compare(got, {c: v for c, v in fd.items() if c != "polarity"}, sites, tol, precision)
    compare(got, {"polarity": fd["polarity"]}, [3], tol, precision)


@pytest.mark.parametrize("precision", ["double", "mixed"])
def test_position_gradient_is_minus_conservative_forces(precision):
    print(f"[gpu] gradients vs forces water216x {precision}", flush=True)
    force, pos, box = make_waterbox("water216x", target_eps=1e-8)
    k = _kernel(force, box, precision)
    got = k.energyGradients(pos, box)
    kc = HipCalcMPIDForceKernel(precision, conservative_torques=True)
    kc.initialize(force.getNumMultipoles(), force, default_box=box)
    e, f = kc.execute(pos, box)
    err = np.sqrt(((got["positions"] + f) ** 2).sum() / (f ** 2).sum())
    print(f"  rel RMS |dE/dr + F| {err:.2e}, energy {got['energy']:.9f} vs {e:.9f}", flush=True)
    # the gradient's dipoles are refined beyond the evaluation's 1e-8 Debye
    assert err <= (1e-7 if precision == "double" else 1e-5)
    assert abs(got["energy"] - e) <= 1e-8 * abs(e)


@pytest.mark.parametrize("precision", ["double", "mixed"])
def test_triclinic_cell(precision):
    """A reduced triclinic cell (32 waters moved molecule by molecule into
    b = (0.31 L, L, 0), c = (-0.42 L, 0.27 L, L)): the reciprocal potential
    vectors through the full grid-coordinate transform.  The cutoff is
    0.35 nm: the cell's smallest height is 0.86 L = 0.85 nm, and a cutoff
    above half of it (the box's own 0.49 nm) would reach two images of some
    pairs, where the single minimum image makes the energy jump."""
    from mpidopenmmplugin_amd.waterbox import ewald_alpha
    from tests.helpers import sheared_waters

    print(f"[gpu] gradients triclinic small box {precision}", flush=True)
    force, pos, box = make_small_waterbox(32, seed=7)
    grid = force.getPMEParameters()[1]
    force.setCutoffDistance(0.35)
    force.setPMEParameters(ewald_alpha(0.35), grid, grid, grid)
    L = box[0]
    vecs = np.array([[L, 0.0, 0.0], [0.31 * L, L, 0.0], [-0.42 * L, 0.27 * L, L]])
    pos = sheared_waters(pos, L, vecs)
    force.setMutualInducedTargetEpsilon(1e-10 if precision == "double" else 1e-6)
    flat = force.flatten()
    sites = [0, 4]  # an oxygen and a hydrogen

    def make():
        fd = fd_gradients(flat, pos, vecs, sites, ["positions", "charges", "dipoles", "quadrupoles", "octopoles"])
        fd.update(fd_gradients(flat, pos, vecs, [3], ["polarity"]))
        return fd
    fd = cached("triclinic", make)
    got = _kernel(force, vecs, precision).energyGradients(pos, vecs, all_sites=True)
    tol = 1e-6 if precision == "double" else 1e-4
    compare(got, {c: v for c, v in fd.items() if c != "polarity"}, sites, tol, precision)
    compare(got, {"polarity": fd["polarity"]}, [3], tol, precision)


def test_extrapolated_is_unsupported():
    force, pos, box = make_small_waterbox(32, seed=7, polarization=MPIDForce.Extrapolated)
    with pytest.raises(MPIDUnsupportedError):
        _kernel(force, box, "double").energyGradients(pos, box)


@pytest.mark.parametrize("precision", ["double", "mixed"])
def test_water4096_per_type_gradients(precision):
    """The gradient workload: the 12,288-atom box's per-type parameters (two
    charges; O dipole, 5 quadrupole and 7 octopole components;
    3 polarisabilities; Thole factors and default width) through the
    torch.autograd.Function, against the committed central differences of
    the FP64 oracle (tests/golden/make_adjoint_fixture.py), each gradient to
    1e-5 of itself.  The per-site Thole factors do not enter this model's
    energy (the pairs that would use them have an unpolarisable hydrogen):
    both gradients are zero."""
    import os

    import torch

    from mpidopenmmplugin_amd.autograd import MPIDEnergy, expand_water_parameters, water_type_parameters

    print(f"[gpu] gradients water4096 per type {precision}", flush=True)
    ref = np.load(os.path.join(os.path.dirname(__file__), "golden", "water4096_gradients_fd.npz"))
    force, pos, box = make_waterbox("water4096", target_eps=1e-8 if precision == "double" else 1e-6)
    assert abs(pos.sum() - float(ref["positions_checksum"])) < 1e-9
    model = MPIDEnergy(force, box, precision)
    p = water_type_parameters(force, dtype=torch.float64)
    x = torch.tensor(pos, dtype=torch.float64, requires_grad=True)
    energy = model(x, *expand_water_parameters(p, force.getNumMultipoles() // 3), p["default_thole"])
    energy.backward()
    names = ("q_O", "q_H", "d_O", "Q_O", "O_O", "alpha_O", "thole_O", "thole_H", "default_thole")
    got = np.concatenate([np.atleast_1d(p[k].grad.numpy()) for k in names])
    fd = ref["gradient"]
    e0 = float(ref["energy"])
    print(f"  energy {float(energy):.9f} oracle {e0:.9f}; adjoint iterations {model.last['adjoint_iterations']}",
          flush=True)
    assert abs(float(energy) - e0) <= (1e-9 if precision == "double" else 1e-6) * abs(e0)
    zero = fd == 0.0
    assert list(np.nonzero(zero)[0]) == [20, 21]  # the two Thole factors
    assert np.all(np.abs(got[zero]) <= 1e-9)
    rel = np.abs(got[~zero] - fd[~zero]) / np.abs(fd[~zero])
    for name, g, f, r in zip(np.asarray(ref["names"])[~zero], got[~zero], fd[~zero], rel):
        print(f"  {name:14s} {g: .9e} fd {f: .9e} rel {r:.1e}", flush=True)
    assert rel.max() <= 1e-5
    # the position gradient is minus the (conservative) forces
    assert x.grad is not None and torch.isfinite(x.grad).all()
