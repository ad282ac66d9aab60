"""The oracle's ZBisect and ThreeFold local frames, which none of the
reference's golden systems use: the lab-frame dipoles against the frames
written out from their definition (MPIDReferenceForce.cpp:398-470), and the
forces, torques mapped onto the frame atoms included, against central
differences of the energy."""
import numpy as np
import pytest

from mpidopenmmplugin_amd import MPIDForce
from oracle import oracle
from tests.helpers import frames_cluster


def unit(v):
    return v / np.linalg.norm(v)


def frame(axis, pos, i, z, x, y):
    """(x, y, z) lab axes of a site, from the definitions: ZBisect z to the
    z atom and x along the sum of the unit vectors to x and y; ThreeFold z
    along the sum of the unit vectors to z, x and y and x to the x atom;
    x is then made orthogonal to z and y = z cross x."""
    ez = unit(pos[z] - pos[i])
    ex = unit(pos[x] - pos[i])
    if axis == MPIDForce.ZBisect:
        ex = unit(ex + unit(pos[y] - pos[i]))
    elif axis == MPIDForce.ThreeFold:
        ez = unit(ez + ex + unit(pos[y] - pos[i]))
    elif axis == MPIDForce.Bisector:
        ez = unit(ez + ex)
    ex = unit(ex - ez * (ez @ ex))
    return ex, np.cross(ez, ex), ez


@pytest.mark.parametrize("periodic", [False, True], ids=["nocutoff", "pme"])
def test_lab_frame_dipoles_of_every_frame(periodic):
    f, pos, box = frames_cluster(periodic=periodic)
    res = oracle.evaluate(f.flatten(), pos, box, details=True)
    lab = res["lab_dipoles"]
    seen, chirality = set(), set()
    for i in range(f.getNumMultipoles()):
        q, d, quad, octo, axis, z, x, y, *_ = f.getMultipoleParameters(i)
        if axis in (MPIDForce.ZBisect, MPIDForce.ThreeFold, MPIDForce.ZThenX, MPIDForce.Bisector):
            ex, ey, ez = frame(axis, pos, i, z, x, y)
            sy = 1.0
            if axis == MPIDForce.ZThenX and y >= 0:
                # a ZThenX site whose y atom lies on the other side flips its
                # y component (checkChiralCenterAtParticle, :357-380)
                volume = np.cross(pos[z] - pos[y], pos[x] - pos[y]) @ (pos[i] - pos[y])
                sy = -1.0 if volume < 0.0 else 1.0
                chirality.add(sy)
            want = d[0] * ex + sy * d[1] * ey + d[2] * ez
            assert np.allclose(lab[i], want, rtol=0, atol=1e-14), (i, axis)
            seen.add(axis)
    assert {MPIDForce.ZBisect, MPIDForce.ThreeFold} <= seen
    assert chirality == {1.0, -1.0}


def _fd_worst(f, pos, box, atoms):
    flat = f.flatten()
    forces = oracle.evaluate(flat, pos, box)["forces"]
    h = 1e-5
    worst = 0.0
    for atom in atoms:
        for k in range(3):
            p = pos.copy()
            p[atom, k] += h
            ep = oracle.evaluate(flat, p, box)["energy"]
            p[atom, k] -= 2 * h
            em = oracle.evaluate(flat, p, box)["energy"]
            worst = max(worst, abs((em - ep) / (2 * h) - forces[atom, k]) / max(1.0, abs(forces[atom, k])))
    return worst


def test_forces_are_energy_gradients():
    """Central differences (h = 1e-5 nm) on atoms of every frame type but
    ThreeFold (ZBisect, ZThenX with and without a chirality y atom, in both
    handednesses, Bisector, ZOnly, NoAxisType, and the frame atoms that take
    their torques), mutual polarization converged to 1e-10 Debye, no
    cutoff."""
    f, pos, box = frames_cluster(target_eps=1e-10, threefold=False)
    assert _fd_worst(f, pos, box, (0, 1, 2, 3, 5, 7, 10, 11, 13, 15)) <= 1e-4


def test_reference_threefold_torque_is_not_a_gradient():
    """The reference maps a ThreeFold site's torque onto its frame atoms with
    an approximate projection (MPIDReferenceForce.cpp:2068-2097, each term
    divided by 3), which the oracle and the gfx950 kernels reproduce: the
    forces then differ from the energy gradient by far more than the
    differencing error (here several kJ/mol/nm against ~1e-7 without
    ThreeFold sites)."""
    f, pos, box = frames_cluster(target_eps=1e-10, polarization=MPIDForce.Direct)
    assert _fd_worst(f, pos, box, (0,)) > 1e-2


# ---------------------------------------------------------------------------
# The ThreeFold torque mapping (mapTorqueToForceForParticle,
# MPIDReferenceForce.cpp:2068-2097) is an approximate projection whose forces
# are not the energy gradient, so finite differences cannot pin it.  These
# checks pin it by properties that the formula has and that a dropped term, a
# wrong sign, a wrong bond length or a wrong prefactor would break.
# ---------------------------------------------------------------------------
def _threefold(pos_i, legs):
    """A ThreeFold site 0 whose z, x and y atoms sit at pos_i + legs[k]."""
    f = MPIDForce()
    z3, z6, z10 = [0.0] * 3, [0.0] * 6, [0.0] * 10
    f.addMultipole(0.0, z3, z6, z10, MPIDForce.ThreeFold, 1, 2, 3, 0.0, z3)
    for _ in range(3):
        f.addMultipole(0.0, z3, z6, z10, MPIDForce.NoAxisType, -1, -1, -1, 0.0, z3)
    pos = np.vstack([pos_i] + [pos_i + np.asarray(l, dtype=np.float64) for l in legs])
    return f.flatten(), pos


def _map(pos_i, legs, tau):
    flat, pos = _threefold(pos_i, legs)
    t = np.zeros((4, 3))
    t[0] = tau
    return oracle.torque_to_force(flat, pos, t), pos


def test_threefold_forces_are_perpendicular_and_balanced():
    """Each frame atom's force is perpendicular to its bond vector, and the
    four forces sum to zero, for general (non-symmetric) geometries."""
    rng = np.random.default_rng(17)
    for _ in range(20):
        pos_i = rng.normal(size=3)
        legs = rng.normal(size=(3, 3)) * rng.uniform(0.5, 2.0, size=(3, 1))
        F, pos = _map(pos_i, legs, rng.normal(size=3))
        for k in range(3):
            assert abs(F[1 + k] @ legs[k]) <= 1e-12 * np.linalg.norm(F[1 + k]) * np.linalg.norm(legs[k]) + 1e-15
        assert np.abs(F.sum(axis=0)).max() <= 1e-12 * np.abs(F).max()


def test_threefold_orthogonal_tripod_closed_form():
    """Bonds along x, y and z with lengths a, b, c: the formula reduces by
    hand (every sine is 1, the cross products are unit axes) to
    F_z-atom = -(y (tx - tz) + z (ty - tx)) / 3a,
    F_x-atom = -(x (tz - ty) + z (ty - tx)) / 3b,
    F_y-atom = -(x (tz - ty) + y (tx - tz)) / 3c,
    and the centre takes minus their sum."""
    a, b, c = 0.7, 1.3, 2.1
    tau = np.array([0.37, -1.1, 0.52])
    tx, ty, tz = tau
    F, _ = _map(np.zeros(3), [[a, 0, 0], [0, b, 0], [0, 0, c]], tau)
    want_u = -np.array([0.0, tx - tz, ty - tx]) / (3 * a)
    want_v = -np.array([tz - ty, 0.0, ty - tx]) / (3 * b)
    want_w = -np.array([tz - ty, tx - tz, 0.0]) / (3 * c)
    assert np.allclose(F[1], want_u, rtol=0, atol=1e-14)
    assert np.allclose(F[2], want_v, rtol=0, atol=1e-14)
    assert np.allclose(F[3], want_w, rtol=0, atol=1e-14)
    assert np.allclose(F[0], -(want_u + want_v + want_w), rtol=0, atol=1e-14)


@pytest.mark.parametrize("tilt", [-0.6, -0.33, 0.0, 0.4])
def test_threefold_symmetric_tripod_torque(tilt):
    """On a symmetric tripod (three equal legs at 120 degrees about an axis)
    the mapped forces exert exactly the part of the torque perpendicular to
    the three-fold axis, and a torque along that axis maps to no force at
    all: rotating the site about its own symmetry axis is what the
    projection leaves out."""
    axis = np.array([0.2, -0.5, 0.84])
    axis /= np.linalg.norm(axis)
    e1 = np.cross(axis, [1.0, 0.0, 0.0])
    e1 /= np.linalg.norm(e1)
    e2 = np.cross(axis, e1)
    legs = [np.cos(p) * e1 + np.sin(p) * e2 - tilt * axis for p in (0.3, 0.3 + 2 * np.pi / 3, 0.3 + 4 * np.pi / 3)]
    pos_i = np.array([0.4, 1.0, -0.2])
    F, _ = _map(pos_i, legs, axis * 1.7)
    assert np.abs(F).max() <= 1e-14
    rng = np.random.default_rng(23)
    for _ in range(5):
        tau = rng.normal(size=3)
        F, pos = _map(pos_i, legs, tau)
        net = sum(np.cross(pos[1 + k] - pos_i, F[1 + k]) for k in range(3))
        assert np.allclose(net, tau - (tau @ axis) * axis, rtol=0, atol=1e-13)


# ---------------------------------------------------------------------------
# With conservative torques (flag 1) a ThreeFold site's torque is mapped by the
# exact chain rule of its frame, so the forces are the energy gradient.
# ---------------------------------------------------------------------------
def _fd_worst_conservative(f, pos, box, atoms):
    flat = f.flatten()
    forces = oracle.evaluate(flat, pos, box, conservative_torques=True)["forces"]
    h = 1e-5
    worst = 0.0
    for atom in atoms:
        for k in range(3):
            p = pos.copy()
            p[atom, k] += h
            ep = oracle.evaluate(flat, p, box, conservative_torques=True)["energy"]
            p[atom, k] -= 2 * h
            em = oracle.evaluate(flat, p, box, conservative_torques=True)["energy"]
            worst = max(worst, abs((em - ep) / (2 * h) - forces[atom, k]) / max(1.0, abs(forces[atom, k])))
    return worst


@pytest.mark.parametrize("polarization", [MPIDForce.Direct, MPIDForce.Mutual])
def test_conservative_threefold_forces_are_energy_gradients(polarization):
    """The ThreeFold centres and their three frame atoms of two molecules
    (one of them a mirror image), central differences h = 1e-5 nm: the same
    agreement as every other frame type, where the reference's projection
    misses by several kJ/mol/nm."""
    f, pos, box = frames_cluster(target_eps=1e-10, polarization=polarization)
    assert _fd_worst_conservative(f, pos, box, (0, 1, 2, 3, 8, 9)) <= 1e-4


def test_conservative_threefold_mapping_invariants():
    """The exact mapping of a torque on a ThreeFold site: the four forces sum
    to zero and their net moment about the site is the torque itself (every
    component, including the one along the three-fold axis that the
    reference's projection drops)."""
    rng = np.random.default_rng(29)
    for _ in range(20):
        pos_i = rng.normal(size=3)
        legs = rng.normal(size=(3, 3)) * rng.uniform(0.5, 2.0, size=(3, 1))
        flat, pos = _threefold(pos_i, legs)
        tau = rng.normal(size=3)
        t = np.zeros((4, 3))
        t[0] = tau
        F = oracle.torque_to_force(flat, pos, t, conservative_torques=True)
        assert np.abs(F.sum(axis=0)).max() <= 1e-12 * np.abs(F).max()
        net = sum(np.cross(pos[1 + k] - pos_i, F[1 + k]) for k in range(3))
        assert np.allclose(net, tau, rtol=0, atol=1e-12 * max(1.0, np.abs(F).max()))
