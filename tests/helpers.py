"""Shared helpers for the test suite: build MPIDForce objects from the golden
fixtures and compare results with the reference's tolerance rule."""
import json
import os

import numpy as np

from mpidopenmmplugin_amd.mpidforce import MPIDForce

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load_reference_tests():
    with open(os.path.join(GOLDEN, "reference_tests.json")) as f:
        return json.load(f)


def force_from_case(case, systems):
    """MPIDForce + positions + box for one golden case (mirrors the set-up code of
    TestReferenceMPIDForce.cpp: make_* then the per-test setters)."""
    sysd = systems[case["system"]]
    f = MPIDForce()
    for a, atom in enumerate(sysd["atoms"]):
        f.addMultipole(atom["charge"], atom["dipole"], atom["quadrupole"], atom["octopole"],
                       atom["axisType"], atom["atomZ"], atom["atomX"], atom["atomY"],
                       atom["thole"], atom["alphas"])
        if atom["cov12"]:
            f.setCovalentMap(a, MPIDForce.Covalent12, atom["cov12"])
        if atom["cov13"]:
            f.setCovalentMap(a, MPIDForce.Covalent13, atom["cov13"])
        if atom.get("cov14"):
            f.setCovalentMap(a, MPIDForce.Covalent14, atom["cov14"])
    f.setNonbondedMethod(case["method"])
    f.setPMEParameters(case["alpha"], case["grid"], case["grid"], case["grid"])
    f.setCutoffDistance(case["cutoff"])
    f.setDefaultTholeWidth(case["default_thole"])
    f.setPolarizationType(case["polarization"])
    f.setMutualInducedTargetEpsilon(case["target_eps"])
    if "scale14" in case:
        f.set14ScaleFactor(case["scale14"])
    pos = np.array(sysd["positions"], dtype=np.float64)
    box = np.array([case["box"]] * 3, dtype=np.float64)
    return f, pos, box


def assert_equal_tol(expected, found, tol):
    """OpenMM's ASSERT_EQUAL_TOL: |e-f| / max(|e|, 1) <= tol."""
    expected = np.asarray(expected, dtype=np.float64)
    found = np.asarray(found, dtype=np.float64)
    scale = np.maximum(np.abs(expected), 1.0)
    err = np.abs(expected - found) / scale
    assert np.all(err <= tol), f"max scaled error {err.max():.3e} > {tol:.1e}\nexpected {expected}\nfound {found}"


def assert_equal_vec(expected, found, tol):
    """OpenMM's ASSERT_EQUAL_VEC: per-vector |e-f| / max(|e|, 1) <= tol (AssertionUtilities.h)."""
    expected = np.asarray(expected, dtype=np.float64).reshape(-1, 3)
    found = np.asarray(found, dtype=np.float64).reshape(-1, 3)
    for e, f in zip(expected, found):
        norm = np.sqrt(e @ e)
        diff = np.sqrt((e - f) @ (e - f))
        assert diff / max(norm, 1.0) <= tol, f"vector mismatch {e} vs {f}: {diff / max(norm, 1.0):.3e}"
