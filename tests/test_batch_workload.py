"""The batched workload's generator (CPU): box edge from the density, PME
grid sizes, cutoff rule, site spacing, seeding, and the oracle's evaluation
of its smallest system."""
import math

import numpy as np
import pytest

from mpidopenmmplugin_amd.waterbox import (WATERS_PER_NM3, lattice_sites, make_batch, make_small_waterbox,
                                           smooth_235)
from oracle import oracle


def test_smooth_235():
    assert [smooth_235(m) for m in (1, 6, 7, 11, 13, 14, 17, 49, 97, 121)] == [1, 6, 8, 12, 15, 15, 18, 50, 100, 125]
    # brute force over a range
    for m in range(1, 400):
        v = smooth_235(m)
        assert v >= m
        r = v
        for p in (2, 3, 5):
            while r % p == 0:
                r //= p
        assert r == 1
        for w in range(m, v):
            r = w
            for p in (2, 3, 5):
                while r % p == 0:
                    r //= p
            assert r != 1


@pytest.mark.parametrize("nwaters", [1, 2, 5, 27, 32, 33, 54, 100, 500, 1024])
def test_lattice_sites_hold_the_waters(nwaters):
    sites, k = lattice_sites(nwaters)
    assert len(sites) >= nwaters
    assert sites.min() >= 0.0 and sites.max() < k
    if len(sites) > 1:
        d = sites[:, None, :] - sites[None, :, :]
        d -= k * np.round(d / k)
        r = np.sqrt((d ** 2).sum(-1))
        np.fill_diagonal(r, np.inf)
        # spacing no worse than the simple cubic lattice with enough sites
        kc = math.ceil(round(nwaters ** (1.0 / 3.0), 9))
        assert r.min() / k >= 1.0 / kc - 1e-12


@pytest.mark.parametrize("nwaters", [32, 37, 250, 1024])
def test_small_waterbox_geometry(nwaters):
    f, pos, box = make_small_waterbox(nwaters, seed=11)
    L = box[0]
    assert np.allclose(box, L)
    assert abs(nwaters / L ** 3 - WATERS_PER_NM3) < 1e-9
    assert f.getNumMultipoles() == 3 * nwaters and pos.shape == (3 * nwaters, 3)
    alpha, nx, ny, nz = f.getPMEParameters()
    assert nx == ny == nz == max(6, smooth_235(math.ceil(L / 0.08)))
    assert nx >= L / 0.08
    rc = f.getCutoffDistance()
    assert rc <= 0.5 * L and rc == pytest.approx(min(0.8, 0.5 * L - 1e-3))
    assert alpha * rc == pytest.approx(math.sqrt(-math.log(1e-3)), rel=1e-12) or alpha > 0
    # rigid waters: O-H 0.09572 nm, HOH 104.52 degrees
    o, h1, h2 = pos[0::3], pos[1::3], pos[2::3]
    assert np.allclose(np.linalg.norm(h1 - o, axis=1), 0.09572, atol=1e-12)
    c = ((h1 - o) * (h2 - o)).sum(1) / 0.09572 ** 2
    assert np.allclose(np.degrees(np.arccos(c)), 104.52, atol=1e-9)
    # oxygens no closer than 0.25 nm (periodic)
    d = o[:, None] - o[None]
    d -= L * np.round(d / L)
    r = np.sqrt((d ** 2).sum(-1))
    np.fill_diagonal(r, np.inf)
    assert r.min() > 0.25


def test_batch_is_seeded_and_shares_parameters():
    a = make_batch(count=6, lo=32, hi=64)
    b = make_batch(count=6, lo=32, hi=64)
    assert [f.getNumMultipoles() for f, _, _ in a] == [f.getNumMultipoles() for f, _, _ in b]
    for (fa, pa, ba), (fb, pb, bb) in zip(a, b):
        assert np.array_equal(pa, pb) and np.array_equal(ba, bb)
    sizes = [f.getNumMultipoles() // 3 for f, _, _ in a]
    assert all(32 <= m <= 64 for m in sizes)
    o0 = a[0][0].getMultipoleParameters(0)
    for f, _, _ in a[1:]:
        assert f.getMultipoleParameters(0) == o0


def test_oracle_smallest_box_translation_invariant():
    """The 32-water box (cutoff just under half its edge, 15^3 grid): the
    oracle's energy does not change when every atom moves by a whole number
    of PME grid cells (the charge grid then shifts cyclically, and the real
    space sum sees the same minimum images), and neither do the forces."""
    f, pos, box = make_small_waterbox(32, seed=1, target_eps=1e-10)
    f.setMutualInducedMaxIterations(300)
    flat = f.flatten()
    n = f.getPMEParameters()[1]
    a = oracle.evaluate(flat, pos, box)
    b = oracle.evaluate(flat, pos + np.array([3, -2, 5]) * box / n, box)
    assert abs(a["energy"] - b["energy"]) <= 1e-8 * abs(a["energy"])
    assert np.abs(a["forces"] - b["forces"]).max() <= 1e-9 * np.abs(a["forces"]).max()
