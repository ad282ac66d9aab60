"""Seeded synthetic polarizable water boxes for the benchmark and parity workloads.

Workloads (all cubic, periodic, PME with B-spline order 6):

=========  ======  ========  =====  =======  ===============================================
name       waters  box (nm)  grid   cutoff   multipoles / polarisability on O
=========  ======  ========  =====  =======  ===============================================
water216   216     1.862     24^3   0.8      dipole, quadrupole; isotropic alpha 0.001 nm^3
water4096  4096    4.966     64^3   0.8      + octopole; anisotropic diagonal alpha
water32768 32768   9.932     128^3  0.9      as water4096 (98,304 atoms)
=========  ======  ========  =====  =======  ===============================================

Geometry: oxygens on a k x k x k lattice jittered by U(-0.02, 0.02) nm, each
molecule randomly oriented, O-H 0.09572 nm, H-O-H 104.52 deg.  Charges O -0.8,
H +0.4 e.  Local frames as the reference's water test (TestReferenceMPIDForce.cpp:
125-146): O Bisector (z = H1, x = H2), H ZThenX (z = O, x = the other H);
covalent 1-2/1-3 maps as there (:198-224).  Local multipoles are drawn once per
seed for the O type (dipole N(0, 0.005) e nm per component, random traceless
quadrupole x 1e-4 e nm^2, random traceless octopole x 1e-5 e nm^3); H carries a
charge only.  Thole 1.3, default Thole width 5.0, Ewald alpha
sqrt(-ln(2*5e-4))/cutoff, mutual polarisation to 1e-6 Debye.
"""
from __future__ import annotations

import math

import numpy as np

from .mpidforce import MPIDForce

CONFIGS = {
    "water216": dict(index=0, k=6, box=1.862, grid=24, cutoff=0.8, octopole=False, anisotropic=False),
    "water4096": dict(index=1, k=16, box=4.966, grid=64, cutoff=0.8, octopole=True, anisotropic=True),
    "water32768": dict(index=2, k=32, box=9.932, grid=128, cutoff=0.9, octopole=True, anisotropic=True),
    # small test box with the water4096 parameter set (octopoles, anisotropic alpha)
    "water216x": dict(index=5, k=6, box=1.862, grid=24, cutoff=0.8, octopole=True, anisotropic=True),
}
SEED_BASE = 20261202
EWALD_TOL = 5e-4


def ewald_alpha(cutoff, tol=EWALD_TOL):
    return math.sqrt(-math.log(2.0 * tol)) / cutoff


def _random_rotation(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    a, b, c, d = q
    return np.array([
        [a * a + b * b - c * c - d * d, 2 * (b * c - a * d), 2 * (b * d + a * c)],
        [2 * (b * c + a * d), a * a - b * b + c * c - d * d, 2 * (c * d - a * b)],
        [2 * (b * d - a * c), 2 * (c * d + a * b), a * a - b * b - c * c + d * d]])


def random_traceless_quadrupole(rng, scale):
    """Input order XX XY YY XZ YZ ZZ."""
    a = rng.normal(size=(3, 3))
    a = 0.5 * (a + a.T)
    a -= np.trace(a) / 3.0 * np.eye(3)
    a *= scale
    return [a[0, 0], a[0, 1], a[1, 1], a[0, 2], a[1, 2], a[2, 2]]


def random_traceless_octopole(rng, scale):
    """Fully symmetric, traceless in every index pair.  Input order
    XXX XXY XYY YYY XXZ XYZ YYZ XZZ YZZ ZZZ."""
    t = np.zeros((3, 3, 3))
    vals = rng.normal(size=10)
    combos = [(0, 0, 0), (0, 0, 1), (0, 1, 1), (1, 1, 1), (0, 0, 2), (0, 1, 2), (1, 1, 2), (0, 2, 2), (1, 2, 2),
              (2, 2, 2)]
    import itertools
    for v, c in zip(vals, combos):
        for p in set(itertools.permutations(c)):
            t[p] = v
    v = np.einsum("iik->k", t)
    d = np.eye(3)
    o = t - 0.2 * (np.einsum("ij,k->ijk", d, v) + np.einsum("ik,j->ijk", d, v) + np.einsum("jk,i->ijk", d, v))
    o *= scale
    return [o[c] for c in combos]


def _o_parameters(rng, octopole, anisotropic):
    """Per-type O multipoles and polarisability, drawn once per seed."""
    o_dip = list(rng.normal(0.0, 0.005, size=3))
    o_quad = random_traceless_quadrupole(rng, 1e-4)
    o_oct = random_traceless_octopole(rng, 1e-5) if octopole else [0.0] * 10
    o_alpha = list(rng.uniform(0.0008, 0.0012, size=3)) if anisotropic else [0.001] * 3
    return o_dip, o_quad, o_oct, o_alpha


def _waters(rng, sites, spacing):
    """Positions of waters whose oxygens sit at the given lattice sites
    (jittered by U(-0.02, 0.02) nm), each randomly oriented."""
    half = math.radians(104.52) / 2.0
    roh = 0.09572
    body = np.array([[0.0, 0.0, 0.0],
                     [roh * math.sin(half), 0.0, roh * math.cos(half)],
                     [-roh * math.sin(half), 0.0, roh * math.cos(half)]])
    pos = np.zeros((3 * len(sites), 3))
    for w, site in enumerate(sites):
        o = (np.asarray(site, dtype=np.float64) + 0.5) * spacing + rng.uniform(-0.02, 0.02, size=3)
        R = _random_rotation(rng)
        pos[3 * w:3 * w + 3] = o + body @ R.T
    return pos


def _water_force(nw, params):
    o_dip, o_quad, o_oct, o_alpha = params
    zero3, zero6, zero10 = [0.0] * 3, [0.0] * 6, [0.0] * 10
    force = MPIDForce()
    for m in range(nw):
        a = 3 * m
        force.addMultipole(-0.8, o_dip, o_quad, o_oct, MPIDForce.Bisector, a + 1, a + 2, -1, 1.3, o_alpha)
        force.addMultipole(0.4, zero3, zero6, zero10, MPIDForce.ZThenX, a, a + 2, -1, 1.3, zero3)
        force.addMultipole(0.4, zero3, zero6, zero10, MPIDForce.ZThenX, a, a + 1, -1, 1.3, zero3)
        force.setCovalentMap(a, MPIDForce.Covalent12, [a + 1, a + 2])
        force.setCovalentMap(a + 1, MPIDForce.Covalent12, [a])
        force.setCovalentMap(a + 1, MPIDForce.Covalent13, [a + 2])
        force.setCovalentMap(a + 2, MPIDForce.Covalent12, [a])
        force.setCovalentMap(a + 2, MPIDForce.Covalent13, [a + 1])
    force.setDefaultTholeWidth(5.0)
    force.setMutualInducedMaxIterations(60)
    return force


def make_waterbox(name="water216", seed=None, polarization=MPIDForce.Mutual, target_eps=1e-6):
    """(MPIDForce, positions [n,3] nm, box lengths [3] nm) for a named workload."""
    cfg = CONFIGS[name]
    rng = np.random.default_rng(SEED_BASE + cfg["index"] if seed is None else seed)
    k = cfg["k"]
    L = cfg["box"]
    params = _o_parameters(rng, cfg["octopole"], cfg["anisotropic"])
    sites = [(ix, iy, iz) for ix in range(k) for iy in range(k) for iz in range(k)]
    pos = _waters(rng, sites, L / k)
    force = _water_force(k ** 3, params)
    force.setNonbondedMethod(MPIDForce.PME)
    force.setCutoffDistance(cfg["cutoff"])
    force.setPMEParameters(ewald_alpha(cfg["cutoff"]), cfg["grid"], cfg["grid"], cfg["grid"])
    force.setPolarizationType(polarization)
    force.setMutualInducedTargetEpsilon(target_eps)
    return force, pos, np.array([L, L, L])


WATERS_PER_NM3 = 33.428  # 1.0 g/cm^3


def smooth_235(m):
    """The smallest integer >= m whose prime factors are 2, 3 and 5 only."""
    m = max(int(m), 1)
    while True:
        v = m
        for p in (2, 3, 5):
            while v % p == 0:
                v //= p
        if v == 1:
            return m
        m += 1


def lattice_sites(nwaters):
    """Sites for nwaters oxygens in a unit cube: of the simple, body-centred
    and face-centred cubic lattices with the fewest cells that hold nwaters
    sites, the one with the widest nearest-neighbour spacing (k**3, 2 k**3
    and 4 k**3 sites; spacing 1/k, sqrt(3)/2k and 1/(sqrt(2) k)).  Returns
    (sites [m,3] in lattice units, spacing of the unit cube in lattice
    units)."""
    best = None
    for basis, nn in (([(0, 0, 0)], 1.0),
                      ([(0, 0, 0), (0.5, 0.5, 0.5)], math.sqrt(3.0) / 2.0),
                      ([(0, 0, 0), (0.5, 0.5, 0), (0.5, 0, 0.5), (0, 0.5, 0.5)], 1.0 / math.sqrt(2.0))):
        k = 1
        while len(basis) * k ** 3 < nwaters:
            k += 1
        if best is None or nn / k > best[0]:
            sites = [(ix + b[0], iy + b[1], iz + b[2]) for ix in range(k) for iy in range(k) for iz in range(k)
                     for b in basis]
            best = (nn / k, sites, k)
    _, sites, k = best
    return np.array(sites, dtype=np.float64), k


def make_small_waterbox(nwaters, seed, params=None, polarization=MPIDForce.Mutual, target_eps=1e-6):
    """One system of the batched workload: nwaters waters at 1.0 g/cm^3 in a
    cubic periodic box of edge (n / 33.428)^(1/3) nm; PME grid spacing about
    0.08 nm (rounded up to a 2,3,5-smooth size, at least 6); cutoff
    min(0.8, L/2 - 1e-3) so that it never exceeds half the box; Ewald alpha
    sqrt(-ln 1e-3) / cutoff.  The oxygens take a seeded choice of nwaters
    sites of the widest-spaced cubic lattice that holds them (lattice_sites).
    params: the per-type O parameters (the water4096 set: octopoles,
    anisotropic alpha), shared by the systems of a batch."""
    rng = np.random.default_rng(seed)
    if params is None:
        params = _o_parameters(rng, True, True)
    L = (nwaters / WATERS_PER_NM3) ** (1.0 / 3.0)
    sites, k = lattice_sites(nwaters)
    chosen = np.sort(rng.choice(len(sites), size=nwaters, replace=False))
    pos = _waters(rng, sites[chosen], L / k)
    force = _water_force(nwaters, params)
    cutoff = min(0.8, 0.5 * L - 1e-3)
    grid = max(6, smooth_235(math.ceil(L / 0.08)))
    force.setNonbondedMethod(MPIDForce.PME)
    force.setCutoffDistance(cutoff)
    force.setPMEParameters(ewald_alpha(cutoff), grid, grid, grid)
    force.setPolarizationType(polarization)
    force.setMutualInducedTargetEpsilon(target_eps)
    return force, pos, np.array([L, L, L])


def make_batch(count=512, seed=SEED_BASE + 3, lo=32, hi=1024, polarization=MPIDForce.Mutual, target_eps=1e-6):
    """The batched workload: count systems of U(lo, hi) waters (seeded), with
    one per-type parameter set for the batch.  A list of (force, positions,
    box)."""
    rng = np.random.default_rng(seed)
    params = _o_parameters(rng, True, True)
    sizes = rng.integers(lo, hi + 1, size=count)
    return [make_small_waterbox(int(m), seed * 1000 + i, params, polarization, target_eps)
            for i, m in enumerate(sizes)]
