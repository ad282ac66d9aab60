"""The oracle's triclinic PME (reciprocal box, fractional multipoles and
potentials, the sheared influence function and minimum image;
MPIDReferenceForce.cpp:2615-2676, 3048-3267, 3329-3368), pinned by the
lattice: a reduced triclinic cell whose lattice is a sublattice of the
cubic one (vectors (2L,0,0), (L,2L,0), (L,L,L), every off-diagonal at its
largest allowed value) filled with four copies of a cubic water cell is the
same infinite system, so its energy is four times the cubic one and every
copy feels the cubic forces.  The real-space pairs within the cutoff are the
same in both, so only the reciprocal part's discretisation differs; at
alpha = 2.0 and the same grid points per unit length (32 per cubic edge)
that leaves 5e-7 in energy and 5e-6 in force RMS, and the tolerances below
are 4x that.  Then the images: moving whole copies by the lattice vectors b
and c, whose x (and y) components a per-axis wrap would miss, changes
nothing beyond rounding."""
import numpy as np

from mpidopenmmplugin_amd.waterbox import make_waterbox
from oracle import oracle
from tests.helpers import replicate_force, triclinic_supercell

ALPHA, GRID = 2.0, 32


def _rel_rms(a, b):
    return np.sqrt(((a - b) ** 2).sum() / (b ** 2).sum())


def test_triclinic_supercell_is_four_cubic_cells():
    f, pos, box = make_waterbox("water216x", target_eps=1e-9)
    L = box[0]
    n = len(pos)
    fc = replicate_force(f, 1, ALPHA, (GRID, GRID, GRID))
    cubic = oracle.evaluate(fc.flatten(), pos, box)
    f4, p4, vecs = triclinic_supercell(f, pos, L, ALPHA, GRID)
    assert vecs[1, 0] != 0 and vecs[2, 0] != 0 and vecs[2, 1] != 0
    tri = oracle.evaluate(f4.flatten(), p4, vecs)
    e_rel = abs(tri["energy"] / 4 - cubic["energy"]) / abs(cubic["energy"])
    assert e_rel < 2e-6, e_rel
    for c in range(4):
        err = _rel_rms(tri["forces"][c * n:(c + 1) * n], cubic["forces"])
        assert err < 2e-5, (c, err)
    # whole copies moved by lattice vectors: the same system
    moved = p4.copy()
    moved[n:2 * n] += vecs[1]
    moved[2 * n:3 * n] -= vecs[2]
    moved[3 * n:] += vecs[1] + vecs[2]
    same = oracle.evaluate(f4.flatten(), moved, vecs)
    assert abs(same["energy"] - tri["energy"]) <= 1e-11 * abs(tri["energy"])
    assert _rel_rms(same["forces"], tri["forces"]) <= 1e-10
