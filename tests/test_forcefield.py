"""Force-field XML to MPIDForce (the reference's MPIDGenerator), the PDB
reader it uses, and the PME parameters taken from the Ewald tolerance.  CPU
only: parameters and maps, no evaluation."""
import math
import os
import warnings

import numpy as np
import pytest

from mpidopenmmplugin_amd import MPIDForce, XmlSerializer
from mpidopenmmplugin_amd.forcefield import (LJPME, PME, CutoffPeriodic, ForceField, NoCutoff, PDBFile, Topology,
                                             set_axis_type, water_topology)
from mpidopenmmplugin_amd.mpidforce import automatic_pme_parameters, legal_fft_size

FF = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "forcefields")


def ff(name):
    return os.path.join(FF, name)


@pytest.mark.parametrize("k,axis,stripped", [
    ([], MPIDForce.NoAxisType, ["", "", ""]),
    (["A"], MPIDForce.ZOnly, ["A", "", ""]),
    (["-A"], MPIDForce.Bisector, ["A", "", ""]),
    (["A", "B"], MPIDForce.ZThenX, ["A", "B", ""]),
    (["A", "B", "C"], MPIDForce.ZThenX, ["A", "B", "C"]),
    (["-A", "-B"], MPIDForce.Bisector, ["A", "B", ""]),
    (["A", "-B"], MPIDForce.Bisector, ["A", "B", ""]),
    (["-A", "B"], MPIDForce.Bisector, ["A", "B", ""]),
    (["A", "-B", "-C"], MPIDForce.ZBisect, ["A", "B", "C"]),
    (["-A", "-B", "-C"], MPIDForce.ThreeFold, ["A", "B", "C"]),
    (["A", "B", "-C"], MPIDForce.ZThenX, ["A", "B", "C"]),
])
def test_axis_type_from_signs(k, axis, stripped):
    """The rules of the reference's setAxisType, case by case: no z atom, no
    frame; z only; any '-' on z or x makes a bisector; '-' on x and y a
    z-bisector; '-' on all three a three-fold axis."""
    idx = ["T"] + list(k)
    assert set_axis_type(idx) == axis
    assert idx == ["T"] + stripped


@pytest.mark.parametrize("order", [("O", "H1", "H2"), ("H1", "H2", "O")])
def test_swm6_water_frames_and_maps(order):
    top = water_topology(4, order=order)
    system = ForceField(ff("swm6.xml")).createSystem(top, PME, 0.8, defaultTholeWidth=8)
    f = system.getForce(0)
    assert f.getNumMultipoles() == 12 and system.getNumParticles() == 12
    for w in range(4):
        a = {name: 3 * w + k for k, name in enumerate(order)}
        o, h1, h2 = a["O"], a["H1"], a["H2"]
        q, d, quad, octo, axis, z, x, y, thole, pol = f.getMultipoleParameters(o)
        assert (q, axis, z, x, y) == (-1.0614, MPIDForce.Bisector, min(h1, h2), max(h1, h2), -1)
        assert d == [0.0, 0.0, -0.023671684]
        assert quad == [0.000150963, 0.0, 0.00008707, 0.0, 0.0, -0.000238034]
        assert octo[4] == 0.000000426 and octo[6] == 0.000000853 and octo[9] == -0.000001279
        assert thole == 8.0 and pol == [0.00088] * 3
        for h, other in ((h1, h2), (h2, h1)):
            q, d, quad, octo, axis, z, x, y, thole, pol = f.getMultipoleParameters(h)
            assert (q, axis, z, x, y, thole, pol) == (0.5307, MPIDForce.ZThenX, o, other, -1, 0.0, [0.0] * 3)
            assert f.getCovalentMap(h, MPIDForce.Covalent12) == [o]
            assert f.getCovalentMap(h, MPIDForce.Covalent13) == [other]
        assert sorted(f.getCovalentMap(o, MPIDForce.Covalent12)) == sorted([h1, h2])
        assert f.getCovalentMap(o, MPIDForce.Covalent13) == []
        assert system.getParticleMass(o) == 15.999 and system.getParticleMass(h1) == 1.008
    # createForce defaults and arguments (mpidplugin.i:730-787)
    assert f.getPolarizationType() == MPIDForce.Extrapolated
    assert f.getNonbondedMethod() == MPIDForce.PME and f.getCutoffDistance() == 0.8
    assert f.getDefaultTholeWidth() == 8.0
    assert f.get14ScaleFactor() == 1.0  # no coulomb14scale on the <MPIDForce> tag


def test_swm4_values():
    f = ForceField(ff("swm4.xml")).createSystem(water_topology(1), NoCutoff, 1.0).getForce(0)
    q, d, quad, octo, *_ , thole, pol = f.getMultipoleParameters(0)
    assert q == -1.11466 and d[2] == -0.026790125 and quad[5] == -0.000214633 and octo[9] == -0.000001032
    assert pol == [0.000978253] * 3 and thole == 8.0
    assert f.getNonbondedMethod() == MPIDForce.NoCutoff


def test_create_system_arguments():
    top = water_topology(2)
    f = ForceField(ff("swm6.xml")).createSystem(
        top, LJPME, 0.9, polarization="Mutual", ewaldErrorTolerance=1e-4, aEwald=3.5, pmeGridDimensions=[24, 25, 27],
        mutualInducedMaxIterations=77, mutualInducedTargetEpsilon=1e-7, coulomb14scale=0.5).getForce(0)
    assert f.getNonbondedMethod() == MPIDForce.PME  # LJPME maps to PME
    assert f.getPolarizationType() == MPIDForce.Mutual
    assert f.getEwaldErrorTolerance() == 1e-4
    assert f.getPMEParameters() == (3.5, 24, 25, 27)
    assert f.getMutualInducedMaxIterations() == 77 and f.getMutualInducedTargetEpsilon() == 1e-7
    assert f.get14ScaleFactor() == 0.5
    f = ForceField(ff("swm6.xml")).createSystem(water_topology(1), PME, 0.9, polarization="direct").getForce(0)
    assert f.getPolarizationType() == MPIDForce.Direct
    with pytest.raises(ValueError, match="invalide polarization type"):
        ForceField(ff("swm6.xml")).createSystem(water_topology(1), PME, 0.9, polarization="scf")
    with pytest.raises(ValueError, match="cutoff method not available"):
        ForceField(ff("swm6.xml")).createSystem(water_topology(1), CutoffPeriodic, 0.9)


def test_ethane_cluster_no_axis_and_14_maps():
    pdb = PDBFile(ff("ethane_cluster.pdb"))
    top = pdb.topology
    assert top.box == [3.5, 3.5, 3.5]
    assert len(top.atoms) == 8 + 3 * 43 and len(pdb.positions) == len(top.atoms)
    with warnings.catch_warnings():
        warnings.simplefilter("error")
        system = ForceField(ff("ethane_water.xml")).createSystem(top, NoCutoff, 1.0)
    f = system.getForce(0)
    assert f.get14ScaleFactor() == 1.0  # coulomb14scale="1.0" on the tag
    # ethane: C1 C2 H H2 H3 H4 H5 H6 with C1-(H,H2,H3), C2-(H4,H5,H6)
    c1, c2 = 0, 1
    h_c1, h_c2 = [2, 3, 4], [5, 6, 7]
    for i in range(8):
        q, *_, axis, z, x, y, thole, pol = f.getMultipoleParameters(i)
        assert (axis, z, x, y) == (MPIDForce.NoAxisType, -1, -1, -1)
        assert q == (-0.27 if i < 2 else 0.09)
        assert pol == ([0.00068] * 3 if i < 2 else [0.0] * 3)
    assert sorted(f.getCovalentMap(c1, MPIDForce.Covalent12)) == sorted(h_c1 + [c2])
    assert sorted(f.getCovalentMap(c1, MPIDForce.Covalent13)) == h_c2
    for h in h_c1:
        assert f.getCovalentMap(h, MPIDForce.Covalent12) == [c1]
        assert sorted(f.getCovalentMap(h, MPIDForce.Covalent13)) == sorted([c2] + [g for g in h_c1 if g != h])
        assert sorted(f.getCovalentMap(h, MPIDForce.Covalent14)) == h_c2
    # waters: charges only, no frames (no kz on the records)
    q, *_, axis, z, x, y, thole, pol = f.getMultipoleParameters(8)
    assert (q, axis, thole, pol) == (-0.834, MPIDForce.NoAxisType, 8.0, [0.00088] * 3)
    assert sorted(f.getCovalentMap(8, MPIDForce.Covalent12)) == [9, 10]


def test_template_errors():
    top = Topology()
    r = top.addResidue("XYZ")
    top.addAtom("N1", "N", r)
    with pytest.raises(ValueError, match="No template found"):
        ForceField(ff("swm6.xml")).createSystem(top, NoCutoff, 1.0)
    # atoms named unlike the template: matched by elements and bond graph
    top = Topology()
    r = top.addResidue("ETHA")
    c1 = top.addAtom("C1", "C", r)
    c2 = top.addAtom("C2", "C", r)
    hs = [top.addAtom(f"H{k}", "H", r) for k in range(6)]
    top.addBond(c1, c2)
    for k, h in enumerate(hs):
        top.addBond(c1 if k < 3 else c2, h)
    f = ForceField(ff("ethane_water.xml")).createSystem(top, NoCutoff, 1.0).getForce(0)
    assert f.getNumMultipoles() == 8


def test_multiple_mpid_tags():
    import xml.etree.ElementTree as ET
    a = ET.fromstring('<ForceField><AtomTypes><Type name="A" class="A" element="O" mass="16"/>'
                      '<Type name="B" class="B" element="H" mass="1"/></AtomTypes>'
                      '<MPIDForce coulomb14scale="0.5"><Multipole type="A" c0="-1"/></MPIDForce></ForceField>')
    b = ET.fromstring('<ForceField><MPIDForce coulomb14scale="0.5"><Multipole type="B" c0="1"/>'
                      '<Polarize type="B" polarizabilityXX="0.001" polarizabilityYY="0.001" '
                      'polarizabilityZZ="0.001" thole="1.3"/></MPIDForce></ForceField>')
    fld = ForceField(a, b)
    gen = fld._generators[0]
    assert set(gen.typeMap) == {"A", "B"} and gen.typeMap["B"][0]["thole"] == 1.3
    c = ET.fromstring('<ForceField><MPIDForce coulomb14scale="0.8"><Multipole type="A" c0="-1"/></MPIDForce>'
                      '</ForceField>')
    with pytest.raises(ValueError, match="different coulomb14scale"):
        ForceField(a, c)
    d = ET.fromstring('<ForceField><MPIDForce><Polarize type="A" polarizabilityXX="0.001" polarizabilityYY="0.001" '
                      'polarizabilityZZ="0.001" thole="1.3"/></MPIDForce></ForceField>')
    types_only = ET.fromstring('<ForceField><AtomTypes><Type name="A" class="A" element="O" mass="16"/>'
                               '</AtomTypes></ForceField>')
    with pytest.raises(ValueError, match="polarize type not present"):
        ForceField(types_only, d)


def test_force_section_before_its_types():
    import xml.etree.ElementTree as ET
    forces = ET.fromstring('<ForceField><MPIDForce><Multipole type="A" c0="-1"/></MPIDForce></ForceField>')
    types = ET.fromstring('<ForceField><AtomTypes><Type name="A" class="A" element="Na" mass="23"/></AtomTypes>'
                          '<Residues><Residue name="ION"><Atom name="NA" type="A"/></Residue></Residues></ForceField>')
    top = Topology()
    top.addAtom("NA", "Na", top.addResidue("ION"))
    f = ForceField(forces, types).createSystem(top, NoCutoff, 1.0).getForce(0)
    assert f.getMultipoleParameters(0)[0] == -1.0 and f.getMultipoleParameters(0)[4] == MPIDForce.NoAxisType


def test_conflicting_file_and_argument_warns():
    with pytest.warns(UserWarning, match="coulomb14scale"):
        f = ForceField(ff("ethane_water.xml")).createSystem(PDBFile(ff("ethane_cluster.pdb")).topology, NoCutoff, 1.0,
                                                            coulomb14scale=0.5).getForce(0)
    assert f.get14ScaleFactor() == 0.5


def test_generated_force_serializes():
    f = ForceField(ff("swm6.xml")).createSystem(water_topology(3), PME, 0.8).getForce(0)
    g = XmlSerializer.deserialize(XmlSerializer.serialize(f))
    for i in range(f.getNumMultipoles()):
        assert g.getMultipoleParameters(i) == f.getMultipoleParameters(i)
        assert g.getCovalentMaps(i) == f.getCovalentMaps(i)


def test_waterbox_fixture_is_liquid_water():
    d = np.load(ff("waterbox_996.npz"))
    pos = d["positions_mA"] * 1e-4
    L = float(d["box_A"]) / 10.0
    assert pos.shape == (2988, 3) and L == pytest.approx(3.1289)
    o, h1 = pos[0::3], pos[1::3]
    assert np.allclose(np.linalg.norm(h1 - o, axis=1), 0.09572, atol=2e-3)
    assert 996 / L ** 3 == pytest.approx(32.5, rel=0.05)  # about 1 g/cm^3


def test_legal_fft_size():
    assert [legal_fft_size(n) for n in (1, 6, 7, 19, 49, 89, 97, 128)] == [6, 6, 8, 20, 50, 90, 100, 128]


@pytest.mark.parametrize("L,rc,raw,legal", [(1.862, 0.8, 19, 20), (4.966, 0.8, 50, 50), (9.932, 0.9, 89, 90)])
def test_automatic_pme_parameters(L, rc, raw, legal):
    """alpha = sqrt(-ln(2 tol)) / r_c and n = ceil(2 alpha L / (3 tol^0.2)):
    at tolerance 5e-4 the three water boxes give alpha 3.2853 / 2.9203 nm^-1
    and grids 19, 50, 89 before rounding to FFT sizes 20, 50, 90."""
    alpha, nx, ny, nz = automatic_pme_parameters(rc, 5e-4, [L, L, L])
    assert alpha == pytest.approx({0.8: 3.2853, 0.9: 2.9203}[rc], abs=1e-4)
    assert math.ceil(2 * alpha * L / (3 * 5e-4 ** 0.2)) == raw
    assert (nx, ny, nz) == (legal, legal, legal)
    # small boxes still get at least 6 points
    assert automatic_pme_parameters(1.0, 5e-4, np.diag([0.5, 0.5, 0.5]))[1:] == (6, 6, 6)  # raw 5
