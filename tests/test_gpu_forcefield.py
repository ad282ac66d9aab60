"""Systems built by the force-field generator, evaluated on the GPU against
the FP64 oracle:

  * SWM6 water (the reference's examples/parameters/swm6.xml) on the
    reference example's 996-water liquid box, PME with a 0.8 nm cutoff and
    alpha and grid left at 0, so both come from the Ewald tolerance (5e-4:
    alpha 3.2853 nm^-1, 32^3 grid); the generator's default extrapolated
    polarization and mutual polarization;
  * the ethane example's force field on ethane with its 43 nearest waters,
    NoCutoff (charge-only sites, no local frames, 1-4 maps across C-C).

Tolerances: extrapolated polarization has no solver, so the DOUBLE path
agrees with the oracle to 1e-8 relative (summation order and fixed-point
force accumulation differ); mutual at the 1e-6 Debye target
agrees within 1e-6 (energy) and 1e-5 (force and induced-dipole relative
RMS) of the oracle converged to 1e-10; the MIXED path within 1e-6 / 1e-5.
"""
import os

import numpy as np
import pytest

from mpidopenmmplugin_amd.forcefield import PME, NoCutoff, ForceField, PDBFile, water_topology
from mpidopenmmplugin_amd.kernel import MPIDContext
from mpidopenmmplugin_amd.mpidforce import automatic_pme_parameters
from oracle import oracle

pytestmark = pytest.mark.gpu

FF = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "forcefields")


def rel_rms(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return np.sqrt(((a - b) ** 2).sum() / (b ** 2).sum())


def swm6_box(polarization, **args):
    d = np.load(os.path.join(FF, "waterbox_996.npz"))
    pos = d["positions_mA"] * 1e-4
    L = float(d["box_A"]) / 10.0
    system = ForceField(os.path.join(FF, "swm6.xml")).createSystem(
        water_topology(996), PME, 0.8, defaultTholeWidth=8, polarization=polarization, **args)
    return system.getForce(0), pos, np.array([L, L, L])


@pytest.fixture(scope="module", params=["extrapolated", "mutual"])
def swm6(request):
    pol = request.param
    force, pos, box = swm6_box(pol, mutualInducedTargetEpsilon=1e-6)
    tight, _, _ = swm6_box(pol, mutualInducedTargetEpsilon=1e-10, mutualInducedMaxIterations=200)
    flat = tight.flatten()
    a, nx, ny, nz = automatic_pme_parameters(0.8, force.getEwaldErrorTolerance(), box)
    flat["alpha"], flat["grid"] = a, (nx, ny, nz)
    print(f"[gpu] swm6 waterbox oracle {pol}", flush=True)
    return pol, force, pos, box, oracle.evaluate(flat, pos, box, details=True), (a, nx, ny, nz)


@pytest.mark.parametrize("precision", ["double", "mixed"])
def test_swm6_waterbox(swm6, precision):
    pol, force, pos, box, ref, pme = swm6
    print(f"[gpu] swm6 waterbox {pol} {precision}", flush=True)
    ctx = MPIDContext(force, box, precision=precision)
    assert ctx.getPMEParametersInContext() == pytest.approx(pme)
    ctx.setPositions(pos)
    st = ctx.getState()
    e, F = st.getPotentialEnergy(), st.getForces()
    mu = ctx.getInducedDipoles()
    exact = pol == "extrapolated" and precision == "double"
    e_tol, f_tol = (1e-8, 1e-8) if exact else (1e-6, 1e-5)
    assert abs(e - ref["energy"]) <= e_tol * abs(ref["energy"])
    assert rel_rms(F, ref["forces"]) <= f_tol
    assert rel_rms(mu, ref["induced"]) <= f_tol


@pytest.mark.parametrize("precision", ["double", "mixed"])
def test_ethane_cluster_nocutoff(precision):
    pdb = PDBFile(os.path.join(FF, "ethane_cluster.pdb"))
    force = ForceField(os.path.join(FF, "ethane_water.xml")).createSystem(pdb.topology, NoCutoff, 1.0).getForce(0)
    pos = np.array(pdb.positions)
    ref = oracle.evaluate(force.flatten(), pos, None, details=True)
    print(f"[gpu] ethane cluster {precision}", flush=True)
    ctx = MPIDContext(force, None, precision=precision)
    ctx.setPositions(pos)
    st = ctx.getState()
    e_tol, f_tol = (1e-8, 1e-8) if precision == "double" else (1e-6, 1e-5)
    assert abs(st.getPotentialEnergy() - ref["energy"]) <= e_tol * abs(ref["energy"])
    assert rel_rms(st.getForces(), ref["forces"]) <= f_tol
    assert rel_rms(ctx.getInducedDipoles(), ref["induced"]) <= f_tol
