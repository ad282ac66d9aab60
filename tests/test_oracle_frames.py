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
