"""XML serialization of MPIDForce in the reference's format, after
serialization/tests/TestSerializeMPIDForce.cpp:59-220 (round trip of every
parameter) plus a document in the form the reference itself writes."""
import random

import numpy as np

from mpidopenmmplugin_amd.mpidforce import MPIDForce
from mpidopenmmplugin_amd.serialization import COVALENT_TYPES, XmlSerializer
from mpidopenmmplugin_amd.waterbox import make_waterbox


def _reference_test_force():
    # the force of TestSerializeMPIDForce.cpp:62-113
    rnd = random.Random(1)
    rand = lambda: float(rnd.randrange(0, 2 ** 31))
    f = MPIDForce()
    f.setForceGroup(3)
    f.setNonbondedMethod(MPIDForce.NoCutoff)
    f.setCutoffDistance(0.9)
    f.setAEwald(0.544)
    f.setPmeGridDimensions([64, 63, 61])
    f.setMutualInducedMaxIterations(200)
    f.setMutualInducedTargetEpsilon(1.0e-05)
    f.setEwaldErrorTolerance(1.0e-05)
    f.set14ScaleFactor(0.4)
    f.setExtrapolationCoefficients([0.0, -0.1, 1.1])
    for ii in range(3):
        f.addMultipole(float(ii + 1), [0.1, rand(), rand()], [rand() for _ in range(6)],
                       [rand() for _ in range(10)], MPIDForce.Bisector, ii + 1, ii + 2, ii + 3, rand(),
                       [rand() for _ in range(3)])
        for jj in range(len(COVALENT_TYPES)):
            f.setCovalentMap(ii, jj, [ii * jj, int(rand()), int(rand())])
    return f


def _assert_same(f1, f2, thole=True):
    assert f1.getForceGroup() == f2.getForceGroup()
    assert f1.getCutoffDistance() == f2.getCutoffDistance()
    assert f1.getNonbondedMethod() == f2.getNonbondedMethod()
    assert f1.getPolarizationType() == f2.getPolarizationType()
    assert f1.getAEwald() == f2.getAEwald()
    assert f1.getMutualInducedMaxIterations() == f2.getMutualInducedMaxIterations()
    assert f1.getMutualInducedTargetEpsilon() == f2.getMutualInducedTargetEpsilon()
    assert f1.getEwaldErrorTolerance() == f2.getEwaldErrorTolerance()
    assert f1.get14ScaleFactor() == f2.get14ScaleFactor()
    assert list(f1.getPmeGridDimensions()) == list(f2.getPmeGridDimensions())
    assert f1.getExtrapolationCoefficients() == f2.getExtrapolationCoefficients()
    if thole:
        assert f1.getDefaultTholeWidth() == f2.getDefaultTholeWidth()
    assert f1.getNumMultipoles() == f2.getNumMultipoles()
    for i in range(f1.getNumMultipoles()):
        assert f1.getMultipoleParameters(i) == f2.getMultipoleParameters(i)
        for t in range(len(COVALENT_TYPES)):
            assert f1.getCovalentMap(i, t) == f2.getCovalentMap(i, t)


def test_round_trip_of_the_reference_test_force():
    f1 = _reference_test_force()
    f2 = XmlSerializer.deserialize(XmlSerializer.serialize(f1))
    _assert_same(f1, f2)


def test_round_trip_keeps_the_evaluated_inputs():
    f1, pos, box = make_waterbox("water216x", polarization=MPIDForce.Extrapolated)
    f2 = XmlSerializer.deserialize(XmlSerializer.serialize(f1))
    _assert_same(f1, f2)
    a, b = f1.flatten(), f2.flatten()
    assert a.keys() == b.keys()
    for k in a:
        assert np.array_equal(np.asarray(a[k]), np.asarray(b[k])), k


def test_reads_the_reference_form():
    """The reference writes an uninitialised ``damp`` and no default Thole
    width (MPIDForceProxy.cpp:70-152); both are ignored on reading, and the
    width takes MPIDForce's default."""
    text = """<?xml version="1.0" ?>
<Force aEwald="3.1" cutoffDistance=".8" ewaldErrorTolerance=".0005" forceGroup="1" mutualInducedMaxIterations="60"
 mutualInducedTargetEpsilon="1e-06" nonbondedMethod="1" polarizationType="2" scaleFactor14="1" type="MPIDForce" version="0">
\t<MultipoleParticleGridDimension d0="24" d1="24" d2="25"/>
\t<ExtrapolationCoefficients c0="-.154" c1=".017" c2=".658" c3=".474"/>
\t<MultipoleParticles>
\t\t<Particle axisType="1" charge="-.8" damp="6.9e-310" multipoleAtomX="2" multipoleAtomY="-1" multipoleAtomZ="1"
 polarizabilityXX=".001" polarizabilityYY=".0011" polarizabilityZZ=".0012" thole="1.3">
\t\t\t<Dipole dX="0" dY="0" dZ=".005"/>
\t\t\t<Quadrupole qXX="1e-4" qXY="0" qXZ="0" qYY="-2e-4" qYZ="0" qZZ="1e-4"/>
\t\t\t<Octopole oXXX="0" oXXY="0" oXXZ="0" oXYY="0" oXYZ="0" oXZZ="0" oYYY="0" oYYZ="0" oYZZ="0" oZZZ="0"/>
\t\t\t<Covalent12><Cv v="1"/><Cv v="2"/></Covalent12>
\t\t\t<Covalent13/><Covalent14/><Covalent15/>
\t\t\t<PolarizationCovalent11><Cv v="1"/><Cv v="2"/></PolarizationCovalent11>
\t\t\t<PolarizationCovalent12/><PolarizationCovalent13/><PolarizationCovalent14/>
\t\t</Particle>
\t</MultipoleParticles>
</Force>
"""
    f = XmlSerializer.deserialize(text)
    assert f.getPMEParameters() == (3.1, 24, 24, 25)
    assert f.getPolarizationType() == MPIDForce.Extrapolated
    assert f.getExtrapolationCoefficients() == [-0.154, 0.017, 0.658, 0.474]
    assert f.getDefaultTholeWidth() == 5.0
    q, d, quad, octo, axis, z, x, y, thole, alphas = f.getMultipoleParameters(0)
    assert (q, axis, z, x, y, thole) == (-0.8, 1, 1, 2, -1, 1.3)
    assert d == [0.0, 0.0, 0.005] and quad == [1e-4, 0.0, -2e-4, 0.0, 0.0, 1e-4] and alphas == [0.001, 0.0011, 0.0012]
    assert f.getCovalentMap(0, 0) == [1, 2] and f.getCovalentMap(0, 4) == [1, 2] and f.getCovalentMap(0, 1) == []
