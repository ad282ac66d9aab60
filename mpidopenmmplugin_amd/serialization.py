"""XML (de)serialization of MPIDForce in the format of the reference's
MPIDForceProxy (serialization/src/MPIDForceProxy.cpp, version 0), as
OpenMM's XmlSerializer writes it: a root element ``Force`` with
``type="MPIDForce"``, properties as attributes, children in creation order.

Differences from the reference, both compatible with its reader
(MPIDForceProxy::deserialize reads neither attribute):
  * ``damp`` carries the damping factor the reference kernels derive from the
    polarizabilities, (mean alpha)^(1/6) (MPIDReferenceKernels.cpp:134); the
    reference writes an uninitialised variable there
    (MPIDForceProxy.cpp:119, 128).
  * ``defaultTholeWidth`` is written as well and read back when present; the
    reference does not serialize it, so a round trip there resets it to the
    MPIDForce default of 5.0.
"""
from __future__ import annotations

import xml.etree.ElementTree as ET

from .mpidforce import MPIDForce

VERSION = 0
# MPIDForceProxy.cpp:44-55, in MPIDForce::CovalentType order
COVALENT_TYPES = ["Covalent12", "Covalent13", "Covalent14", "Covalent15",
                  "PolarizationCovalent11", "PolarizationCovalent12", "PolarizationCovalent13",
                  "PolarizationCovalent14"]
_QUAD = ["qXX", "qXY", "qYY", "qXZ", "qYZ", "qZZ"]
_OCT = ["oXXX", "oXXY", "oXYY", "oYYY", "oXXZ", "oXYZ", "oYYZ", "oXZZ", "oYZZ", "oZZZ"]


def _num(v) -> str:
    return repr(float(v))


def _set(elem, **props):
    for k, v in props.items():
        elem.set(k, v if isinstance(v, str) else (str(int(v)) if isinstance(v, (bool, int)) else _num(v)))


def to_element(force: MPIDForce) -> ET.Element:
    """The serialized form (MPIDForceProxy::serialize, MPIDForceProxy.cpp:70-152)."""
    root = ET.Element("Force")
    alpha, nx, ny, nz = force.getPMEParameters()
    _set(root, type="MPIDForce", version=VERSION, forceGroup=force.getForceGroup(),
         nonbondedMethod=int(force.getNonbondedMethod()), polarizationType=int(force.getPolarizationType()),
         mutualInducedMaxIterations=int(force.getMutualInducedMaxIterations()),
         cutoffDistance=float(force.getCutoffDistance()), aEwald=float(alpha),
         mutualInducedTargetEpsilon=float(force.getMutualInducedTargetEpsilon()),
         ewaldErrorTolerance=float(force.getEwaldErrorTolerance()), scaleFactor14=float(force.get14ScaleFactor()),
         defaultTholeWidth=float(force.getDefaultTholeWidth()))
    _set(ET.SubElement(root, "MultipoleParticleGridDimension"), d0=int(nx), d1=int(ny), d2=int(nz))
    coeffs = ET.SubElement(root, "ExtrapolationCoefficients")
    for i, c in enumerate(force.getExtrapolationCoefficients()):
        coeffs.set(f"c{i}", _num(c))
    particles = ET.SubElement(root, "MultipoleParticles")
    for i in range(force.getNumMultipoles()):
        q, d, quad, octo, axis, z, x, y, thole, alphas = force.getMultipoleParameters(i)
        p = ET.SubElement(particles, "Particle")
        damp = (sum(alphas) / 3.0) ** (1.0 / 6.0) if sum(alphas) > 0.0 else 0.0
        _set(p, axisType=int(axis), multipoleAtomZ=int(z), multipoleAtomX=int(x), multipoleAtomY=int(y),
             charge=float(q), thole=float(thole), damp=float(damp), polarizabilityXX=float(alphas[0]),
             polarizabilityYY=float(alphas[1]), polarizabilityZZ=float(alphas[2]))
        _set(ET.SubElement(p, "Dipole"), dX=float(d[0]), dY=float(d[1]), dZ=float(d[2]))
        qe = ET.SubElement(p, "Quadrupole")
        for name, v in zip(_QUAD, quad):
            qe.set(name, _num(v))
        oe = ET.SubElement(p, "Octopole")
        for name, v in zip(_OCT, octo):
            oe.set(name, _num(v))
        for t, name in enumerate(COVALENT_TYPES):
            m = ET.SubElement(p, name)
            for a in force.getCovalentMap(i, t):
                ET.SubElement(m, "Cv").set("v", str(int(a)))
    return root


def serialize(force: MPIDForce) -> str:
    """XmlSerializer.serialize(force): the XML text."""
    root = to_element(force)
    ET.indent(root, space="\t")
    return '<?xml version="1.0" ?>\n' + ET.tostring(root, encoding="unicode") + "\n"


def from_element(root: ET.Element) -> MPIDForce:
    """MPIDForceProxy::deserialize, MPIDForceProxy.cpp:154-244."""
    version = int(root.get("version"))
    if version < 0 or version > VERSION:
        raise ValueError("Unsupported version number")
    f = MPIDForce()
    f.setForceGroup(int(root.get("forceGroup", "0")))
    f.setNonbondedMethod(int(root.get("nonbondedMethod")))
    f.setPolarizationType(int(root.get("polarizationType")))
    f.setMutualInducedMaxIterations(int(root.get("mutualInducedMaxIterations")))
    f.setCutoffDistance(float(root.get("cutoffDistance")))
    f.setMutualInducedTargetEpsilon(float(root.get("mutualInducedTargetEpsilon")))
    f.setEwaldErrorTolerance(float(root.get("ewaldErrorTolerance")))
    f.set14ScaleFactor(float(root.get("scaleFactor14")))
    if root.get("defaultTholeWidth") is not None:
        f.setDefaultTholeWidth(float(root.get("defaultTholeWidth")))
    g = root.find("MultipoleParticleGridDimension")
    f.setPMEParameters(float(root.get("aEwald")), int(g.get("d0")), int(g.get("d1")), int(g.get("d2")))
    ce = root.find("ExtrapolationCoefficients")
    coeffs = []
    while ce is not None and ce.get(f"c{len(coeffs)}") is not None:
        coeffs.append(float(ce.get(f"c{len(coeffs)}")))
    f.setExtrapolationCoefficients(coeffs)
    for p in root.find("MultipoleParticles"):
        d = p.find("Dipole")
        dip = [float(d.get(k)) for k in ("dX", "dY", "dZ")]
        quad = [float(p.find("Quadrupole").get(k)) for k in _QUAD]
        octo = [float(p.find("Octopole").get(k)) for k in _OCT]
        alphas = [float(p.get(k)) for k in ("polarizabilityXX", "polarizabilityYY", "polarizabilityZZ")]
        i = f.addMultipole(float(p.get("charge")), dip, quad, octo, int(p.get("axisType")),
                           int(p.get("multipoleAtomZ")), int(p.get("multipoleAtomX")), int(p.get("multipoleAtomY")),
                           float(p.get("thole")), alphas)
        for t, name in enumerate(COVALENT_TYPES):
            m = p.find(name)
            f.setCovalentMap(i, t, [int(c.get("v")) for c in (m if m is not None else [])])
    return f


def deserialize(text: str) -> MPIDForce:
    """XmlSerializer.deserialize(text) for an MPIDForce."""
    root = ET.fromstring(text)
    if root.get("type", "MPIDForce") != "MPIDForce":
        raise ValueError("not an MPIDForce: " + str(root.get("type")))
    return from_element(root)


class XmlSerializer:
    """The two XmlSerializer calls the reference's tests make for MPIDForce."""

    @staticmethod
    def serialize(force: MPIDForce) -> str:
        return serialize(force)

    @staticmethod
    def deserialize(text: str) -> MPIDForce:
        return deserialize(text)
