"""Writes tests/golden/water4096_gradients_fd.npz: central finite differences
of the FP64 oracle's energy with respect to the per-type parameters of the
12,288-atom water box (octopoles, anisotropic polarisability, 64^3 grid,
mutual polarisation re-converged to 1e-10 Debye at every point).

Parameters, in this order: q_O, q_H; the O dipole (x, y, z); the O
quadrupole's independent components XX XY YY XZ YZ (ZZ = -XX - YY); the O
octopole's independent components XXY XYY XXZ XYZ YYZ XZZ YZZ (XXX, YYY
and ZZZ follow from the traces); the O polarisabilities (xx, yy, zz); the O
and H Thole factors; the default Thole width.

The differences take the oracle's variational energy (the energy plus
-1/2 eps mu.r, r the residual of the induced dipoles), whose error is second
order in the solver's tolerance: the plain energy moves by ~1e-7 kJ/mol from
one solve to the next, which the small derivatives would not survive.

Steps are relative to each parameter's value, or to the RMS value of its
class (charges, dipole, quadrupole, octopole, polarisabilities, Thole
factors) where that is larger: a component near zero (the quadrupole's YZ
is 4e-9) would otherwise take a step at the rounding level of the energy.
The energy is quadratic in the multipole parameters (the induced dipoles
are linear in them), so their two-point central differences carry no
truncation error and take a step of 1e-2: the rounding of the energy
(~1e-8 kJ/mol) divided by the step is then below 1e-8 of each derivative,
where a step of 1e-4 left 5e-6 on the quadrupole's YZ.  The
polarisabilities and Thole widths take the four-point central stencil
(error O(h^4)) at 1e-4: the Thole damping varies on the scale of
r / damp^2, and the two-point error of the default width's derivative is
~1e-5 at this step.
Each point is one oracle evaluation (about 20-40 s), run in a pool of
processes; the result is committed so that the GPU box only compares.

Run:  python tests/golden/make_adjoint_fixture.py [workers]
"""
import os
import sys
import time
from multiprocessing import Pool

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from mpidopenmmplugin_amd.waterbox import make_waterbox  # noqa: E402
from oracle import oracle  # noqa: E402

QUAD_FREE = (0, 1, 2, 3, 4)       # XX XY YY XZ YZ of the API order XX XY YY XZ YZ ZZ
OCT_FREE = (1, 2, 4, 5, 6, 7, 8)  # XXY XYY XXZ XYZ YYZ XZZ YZZ of XXX XXY XYY YYY XXZ XYZ YYZ XZZ YZZ ZZZ
NAMES = (["q_O", "q_H", "d_O_x", "d_O_y", "d_O_z"] + [f"Q_O_{c}" for c in ("xx", "xy", "yy", "xz", "yz")] +
         [f"O_O_{c}" for c in ("xxy", "xyy", "xxz", "xyz", "yyz", "xzz", "yzz")] +
         ["alpha_O_xx", "alpha_O_yy", "alpha_O_zz", "thole_O", "thole_H", "default_thole"])


def type_vector(flat):
    q = flat["quadrupoles"].reshape(-1, 6)[0]
    o = flat["octopoles"].reshape(-1, 10)[0]
    return np.concatenate([[flat["charges"][0], flat["charges"][1]], flat["dipoles"].reshape(-1, 3)[0],
                           q[list(QUAD_FREE)], o[list(OCT_FREE)], flat["polarity"].reshape(-1, 3)[0],
                           [flat["tholes"][0], flat["tholes"][1], flat["default_thole"]]])


def apply_types(flat, v):
    """flat with the per-type parameter vector v (O, H, H per molecule)."""
    out = dict(flat)
    n = int(flat["n"])
    nw = n // 3
    xx, xy, yy, xz, yz = v[5:10]
    quad = np.array([xx, xy, yy, xz, yz, -(xx + yy)])
    xxy, xyy, xxz, xyz, yyz, xzz, yzz = v[10:17]
    octo = np.array([-(xyy + xzz), xxy, xyy, -(xxy + yzz), xxz, xyz, yyz, xzz, yzz, -(xxz + yyz)])
    charges = np.tile([v[0], v[1], v[1]], nw)
    dip = np.zeros((n, 3))
    dip[0::3] = v[2:5]
    quads = np.zeros((n, 6))
    quads[0::3] = quad
    octs = np.zeros((n, 10))
    octs[0::3] = octo
    pol = np.zeros((n, 3))
    pol[0::3] = v[17:20]
    tholes = np.tile([v[20], v[21], v[21]], nw)
    out.update(charges=charges, dipoles=dip.reshape(-1), quadrupoles=quads.reshape(-1), octopoles=octs.reshape(-1),
               polarity=pol.reshape(-1), tholes=tholes, default_thole=float(v[22]))
    return out


_SYS = None


def _energy(v):
    flat, pos, box = _SYS
    return oracle.evaluate(apply_types(flat, np.asarray(v)), pos, box)["energy_variational"]


def main():
    global _SYS
    workers = int(sys.argv[1]) if len(sys.argv) > 1 else 8
    f, pos, box = make_waterbox("water4096", target_eps=1e-10)
    flat = f.flatten()
    v0 = type_vector(flat)
    assert np.allclose(apply_types(flat, v0)["quadrupoles"], flat["quadrupoles"], atol=1e-15)
    assert np.allclose(apply_types(flat, v0)["octopoles"], flat["octopoles"], atol=1e-15)
    _SYS = (flat, pos, box)
    classes = [slice(0, 2), slice(2, 5), slice(5, 10), slice(10, 17), slice(17, 20), slice(20, 22), slice(22, 23)]
    scale = np.abs(v0).copy()
    for c in classes:
        scale[c] = np.maximum(scale[c], np.sqrt(np.mean(v0[c] ** 2)))
    steps = 1e-4 * scale
    steps[:17] = 1e-2 * scale[:17]  # charges and multipoles: the energy is quadratic in them
    # The default width's derivative (about -2 kJ/mol per unit width) is small
    # against the rounding of a 12,288-atom energy (~1e-8 kJ/mol): at a step
    # of 5e-4 the four-point estimates move by 2e-5 relative from step to
    # step, while at 8e-3 and 3e-2 they agree to 3e-7.  It takes 3e-2.
    steps[22] = 6e-3 * abs(v0[22])
    four = np.zeros(len(v0), dtype=bool)
    four[17:] = True  # polarisabilities, Thole factors, default width
    points = [v0]
    index = []  # (parameter, multiple of the step) per point after the first
    for k in range(len(v0)):
        for m in ((1.0, -1.0, 2.0, -2.0) if four[k] else (1.0, -1.0)):
            v = v0.copy()
            v[k] += m * steps[k]
            points.append(v)
            index.append((k, m))
    t = time.time()
    with Pool(workers) as pool:
        energies = pool.map(_energy, points, chunksize=1)
    e0 = energies[0]
    e = {km: en for km, en in zip(index, energies[1:])}
    ep = np.array([e[(k, 1.0)] for k in range(len(v0))])
    em = np.array([e[(k, -1.0)] for k in range(len(v0))])
    grad = (ep - em) / (2.0 * steps)
    for k in np.nonzero(four)[0]:
        grad[k] = (8.0 * (ep[k] - em[k]) - (e[(k, 2.0)] - e[(k, -2.0)])) / (12.0 * steps[k])
    print(f"{len(points)} oracle evaluations in {time.time() - t:.0f} s; energy {e0:.6f}")
    for name, val, g in zip(NAMES, v0, grad):
        print(f"  {name:14s} {val: .6e}  dE/dp {g: .9e}")
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "water4096_gradients_fd.npz")
    np.savez_compressed(out, names=np.array(NAMES), values=v0, steps=steps, gradient=grad, energy=np.float64(e0),
                        energy_plus=ep, energy_minus=em, positions_checksum=np.float64(pos.sum()))
    print("wrote", out, os.path.getsize(out), "bytes")


if __name__ == "__main__":
    main()
