"""Pins of the oracle's getElectrostaticPotential and getSystemMultipoleMoments
(MPIDReferenceForce.cpp:2352-2531; the reference's tests do not exercise
them) by closed forms and by the golden-pinned pair energy: the potential of
a site, times a unit test charge, is their interaction energy."""
import numpy as np
import pytest

from oracle import oracle
from mpidopenmmplugin_amd.mpidforce import MPIDForce

KE = 138.935455846
DEBYE = 4.80321  # the reference's conversion, MPIDReferenceForce.cpp:2447


def _site(f, q, d=(0, 0, 0), quad=(0, 0, 0, 0, 0, 0), alphas=(0, 0, 0)):
    # NoAxisType: molecular frame = lab frame
    return f.addMultipole(q, list(d), list(quad), [0.0] * 10, MPIDForce.NoAxisType, -1, -1, -1, 0.39, list(alphas))


def _nocutoff(f):
    f.setNonbondedMethod(MPIDForce.NoCutoff)
    f.setPolarizationType(MPIDForce.Mutual)
    f.setMutualInducedTargetEpsilon(1e-12)
    f.setDefaultTholeWidth(0.39)
    return f


def test_potential_of_point_charges_is_coulomb():
    rng = np.random.default_rng(3)
    f = _nocutoff(MPIDForce())
    q = rng.uniform(-1, 1, 5)
    pos = rng.uniform(0, 2, (5, 3))
    for v in q:
        _site(f, v)
    pts = rng.uniform(3, 4, (7, 3))
    phi = oracle.electrostatic_potential(f.flatten(), pos, None, pts)
    want = [KE * sum(q[i] / np.linalg.norm(pos[i] - p) for i in range(5)) for p in pts]
    assert np.allclose(phi, want, rtol=1e-13, atol=0)


def test_potential_matches_pair_energy_with_a_test_charge():
    """A site with a charge, dipole and traceless quadrupole, and a unit
    charge at a point: their NoCutoff interaction energy (the oracle's pair
    formula, pinned by the reference goldens) is the potential times 1."""
    rng = np.random.default_rng(5)
    d = rng.normal(0, 0.02, 3)
    qm = rng.normal(0, 1e-3, (3, 3))
    qm = 0.5 * (qm + qm.T)
    qm -= np.eye(3) * np.trace(qm) / 3
    quad = [qm[0, 0], qm[0, 1], qm[1, 1], qm[0, 2], qm[1, 2], qm[2, 2]]  # API order XX XY YY XZ YZ ZZ
    a = np.array([0.1, 0.2, 0.3])
    for trial in range(4):
        t = a + rng.normal(0, 0.5, 3)
        pair = _nocutoff(MPIDForce())
        _site(pair, -0.7, d, quad)
        _site(pair, 1.0)
        e = oracle.evaluate(pair.flatten(), np.array([a, t]), None)["energy"]
        one = _nocutoff(MPIDForce())
        _site(one, -0.7, d, quad)
        phi = oracle.electrostatic_potential(one.flatten(), np.array([a]), None, t[None, :])[0]
        assert abs(phi - e) <= 1e-12 * abs(e), (trial, phi, e)


def test_potential_adds_the_induced_dipoles():
    """Polarisable sites: the potential equals that of the same sites with the
    converged induced dipoles added to their permanent dipoles and no
    polarisability."""
    rng = np.random.default_rng(7)
    pos = rng.uniform(0, 1.2, (6, 3))
    f = _nocutoff(MPIDForce())
    perm = rng.normal(0, 0.01, (6, 3))
    for i in range(6):
        _site(f, (-1) ** i * 0.4, perm[i], alphas=[1e-3, 1.1e-3, 0.9e-3])
    res = oracle.evaluate(f.flatten(), pos, None, details=True)
    mu = res["induced"].reshape(-1, 3)
    assert np.abs(mu).max() > 1e-4
    g = _nocutoff(MPIDForce())
    for i in range(6):
        _site(g, (-1) ** i * 0.4, perm[i] + mu[i])
    pts = rng.uniform(2.5, 3.5, (5, 3))
    a = oracle.electrostatic_potential(f.flatten(), pos, None, pts)
    b = oracle.electrostatic_potential(g.flatten(), pos, None, pts)
    assert np.allclose(a, b, rtol=1e-11, atol=0)


def test_multipole_moments_closed_forms():
    """Two opposite charges and one point dipole with a quadrupole, no
    polarisability: charge, dipole (e nm -> Debye, x 10 x 4.80321) and the
    traceless quadrupole 0.5 (T - tr/3) + sum of atomic quadrupoles, in
    Buckingham (x 100 x 3 x 4.80321), about the centre of mass."""
    f = _nocutoff(MPIDForce())
    _site(f, 0.5)
    _site(f, -0.5)
    qm = np.array([[2e-4, 1e-4, 0.0], [1e-4, -1e-4, 3e-5], [0.0, 3e-5, -1e-4]])
    _site(f, 0.0, (0.01, -0.02, 0.005), [qm[0, 0], qm[0, 1], qm[1, 1], qm[0, 2], qm[1, 2], qm[2, 2]])
    pos = np.array([[0.3, 0.0, 0.0], [-0.1, 0.2, 0.0], [0.0, 0.0, 0.4]])
    masses = np.array([2.0, 1.0, 3.0])
    out = oracle.system_multipole_moments(f.flatten(), pos, None, masses)
    com = (masses[:, None] * pos).sum(0) / masses.sum()
    x = pos - com
    q = np.array([0.5, -0.5, 0.0])
    d = np.zeros((3, 3))
    d[2] = [0.01, -0.02, 0.005]
    assert abs(out[0]) < 1e-15
    assert np.allclose(out[1:4], ((q[:, None] * x).sum(0) + d.sum(0)) * 10 * DEBYE, rtol=1e-13)
    T = sum(q[i] * np.outer(x[i], x[i]) + np.outer(x[i], d[i]) + np.outer(d[i], x[i]) for i in range(3))
    Q = 0.5 * (T - np.eye(3) * np.trace(T) / 3) + qm
    assert np.allclose(out[4:].reshape(3, 3), Q * 100 * 3 * DEBYE, rtol=1e-12, atol=1e-15)
