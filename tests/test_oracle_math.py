"""Pins of the oracle's building blocks by what the mathematics fixes
(independent of the reference's code): B-spline identities, FFT vs numpy,
symmetry of the dipole interaction operator, Ewald splitting independence."""
import numpy as np
import pytest

from oracle import oracle
from mpidopenmmplugin_amd.mpidforce import MPIDForce
from mpidopenmmplugin_amd.waterbox import make_waterbox


@pytest.mark.parametrize("w", [0.0, 0.13, 0.5, 0.77, 0.999])
def test_bspline_identities(w):
    th = oracle.bspline(w)
    assert abs(th[:, 0].sum() - 1.0) < 1e-14          # partition of unity
    for r in range(1, 5):
        assert abs(th[:, r].sum()) < 1e-13            # derivatives of a constant vanish
    # order-6 cardinal B-spline M6 sampled at w + 5 - k
    def m6(x):
        from math import comb, factorial
        return sum((-1) ** j * comb(6, j) * max(x - j, 0.0) ** 5 for j in range(7)) / factorial(5)
    for k in range(6):
        assert abs(th[k, 0] - m6(w + 5 - k)) < 1e-13
    # derivatives against finite differences of the values
    h = 1e-6
    if 1e-5 < w < 1 - 1e-5:
        thp, thm = oracle.bspline(w + h), oracle.bspline(w - h)
        for r in range(1, 5):
            fd = (thp[:, r - 1] - thm[:, r - 1]) / (2 * h)
            assert np.allclose(fd, th[:, r], atol=1e-6)


@pytest.mark.parametrize("shape", [(6, 8, 10), (24, 24, 24), (15, 12, 7)])
def test_fft_matches_numpy(shape):
    rng = np.random.default_rng(1)
    a = rng.normal(size=shape) + 1j * rng.normal(size=shape)
    assert np.allclose(oracle.fft3d(a, -1), np.fft.fftn(a), atol=1e-9)
    assert np.allclose(oracle.fft3d(a, +1), np.fft.ifftn(a) * a.size, atol=1e-9)


def test_bspline_moduli_no_euler_fix_at_order6():
    # order-6 moduli never fall below the 1e-7 threshold on even grids
    for n in (24, 64, 128):
        m = oracle.bspline_moduli(n)
        assert m.min() > 1e-7


def test_induced_field_operator_is_symmetric():
    f, pos, box = make_waterbox("water216")
    flat = f.flatten()
    rng = np.random.default_rng(7)
    a = rng.normal(size=pos.shape) * 1e-3
    b = rng.normal(size=pos.shape) * 1e-3
    ta = oracle.induced_field(flat, pos, box, a)
    tb = oracle.induced_field(flat, pos, box, b)
    lhs, rhs = (b * ta).sum(), (a * tb).sum()
    assert abs(lhs - rhs) <= 1e-10 * max(abs(lhs), abs(rhs))


def test_energy_independent_of_ewald_alpha(reference_tests):
    """Real + reciprocal + self energy does not depend on the splitting
    parameter once both sums are converged (charges and dipoles, Direct)."""
    from tests.helpers import force_from_case
    case = next(c for c in reference_tests["cases"] if c["name"] == "water_pme_direct")
    energies = []
    for alpha, grid in ((3.2, 64), (3.6, 64)):
        c = dict(case)
        c["alpha"], c["grid"], c["cutoff"] = alpha, grid, 0.99
        f, pos, box = force_from_case(c, reference_tests["systems"])
        energies.append(oracle.evaluate(f.flatten(), pos, box)["energy"])
    assert abs(energies[0] - energies[1]) < 2e-5 * abs(energies[0])
