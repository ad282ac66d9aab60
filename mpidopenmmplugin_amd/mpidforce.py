"""Parameter container mirroring the reference's ``MPIDForce`` API.

The method names, argument meaning, defaults and tuple-returning getters follow
the C++ class (openmmapi/include/openmm/MPIDForce.h:54-489,
openmmapi/src/MPIDForce.cpp:43-277) as exposed by its SWIG module
(python/mpidplugin.i:47-527), so scripts written against ``mpidplugin.MPIDForce``
run unchanged.  Input orderings are those of the reference API:

* dipole ``X Y Z`` (e nm)
* quadrupole ``XX XY YY XZ YZ ZZ`` (e nm^2, one third of the traceless moment)
* octopole ``XXX XXY XYY YYY XXZ XYZ YYZ XZZ YZZ ZZZ`` (e nm^3, one fifteenth)

This file holds parameters only; evaluation happens in :mod:`.kernel` on the GPU.
"""
from __future__ import annotations

import math

import numpy as np


def legal_fft_size(n):
    """The smallest size >= n (and >= 6) whose prime factors are 2, 3 and 5:
    the sizes the gfx950 reciprocal-space FFT transforms."""
    m = max(int(n), 6)
    while True:
        v = m
        for p in (2, 3, 5):
            while v % p == 0:
                v //= p
        if v == 1:
            return m
        m += 1


def automatic_pme_parameters(cutoff, ewald_tolerance, box):
    """Ewald alpha and grid from the error tolerance, as the reference's
    kernels obtain them when alpha or the grid is 0
    (MPIDReferenceKernels.cpp:161-170, MPIDCudaKernels.cpp:441-448): OpenMM's
    NonbondedForceImpl::calcPMEParameters (OpenMM 7/8, not part of the
    reference) sets alpha = sqrt(-ln(2 tol)) / r_c and
    n_k = max(6, ceil(2 alpha L_k / (3 tol^(1/5)))) over the System's default
    box edges L_k.  The CUDA platform then rounds each n_k up to a size its
    FFT handles; here that is the next 2,3,5-smooth size (legal_fft_size).
    box: the three edge lengths (nm) or a 3x3 matrix of box vectors."""
    b = np.asarray(box, dtype=np.float64)
    edges = np.diag(b.reshape(3, 3)) if b.size == 9 else b.reshape(3)
    tol = float(ewald_tolerance)
    alpha = math.sqrt(-math.log(2.0 * tol)) / float(cutoff)
    grid = [legal_fft_size(max(6, math.ceil(2.0 * alpha * L / (3.0 * tol ** 0.2)))) for L in edges]
    return alpha, grid[0], grid[1], grid[2]


class MPIDForceError(RuntimeError):
    """Raised for invalid parameters (the reference throws ``OpenMMException``)."""


class MPIDForce:
    # NonbondedMethod (MPIDForce.h:58-72)
    NoCutoff = 0
    PME = 1
    # PolarizationType (MPIDForce.h:74-97)
    Mutual = 0
    Direct = 1
    Extrapolated = 2
    # MultipoleAxisTypes (MPIDForce.h:99)
    ZThenX = 0
    Bisector = 1
    ZBisect = 2
    ThreeFold = 3
    ZOnly = 4
    NoAxisType = 5
    LastAxisTypeIndex = 6
    # CovalentType (MPIDForce.h:101-103)
    Covalent12 = 0
    Covalent13 = 1
    Covalent14 = 2
    Covalent15 = 3
    PolarizationCovalent11 = 4
    PolarizationCovalent12 = 5
    PolarizationCovalent13 = 6
    PolarizationCovalent14 = 7
    CovalentEnd = 8

    def __init__(self):
        # defaults of MPIDForce::MPIDForce(), MPIDForce.cpp:43-50
        self._nonbonded_method = MPIDForce.NoCutoff
        self._polarization_type = MPIDForce.Extrapolated
        self._cutoff = 1.0
        self._ewald_tol = 5e-4
        self._max_iter = 60
        self._target_eps = 1.0e-5
        self._default_thole = 5.0
        self._alpha = 0.0
        self._grid = [0, 0, 0]
        self._scale14 = 1.0
        self._extrapolation = [-0.154, 0.017, 0.658, 0.474]
        self._force_group = 0
        self._multipoles = []  # list of dicts
        self._covalent = []    # list of [8 lists]

    # ---- global parameters -------------------------------------------------
    def getNumMultipoles(self):
        return len(self._multipoles)

    def getNonbondedMethod(self):
        return self._nonbonded_method

    def setNonbondedMethod(self, method):
        if method not in (MPIDForce.NoCutoff, MPIDForce.PME):
            raise MPIDForceError("MPIDForce: Illegal value for nonbonded method")
        self._nonbonded_method = int(method)

    def getPolarizationType(self):
        return self._polarization_type

    def setPolarizationType(self, ptype):
        self._polarization_type = int(ptype)

    def getExtrapolationCoefficients(self):
        return list(self._extrapolation)

    def setExtrapolationCoefficients(self, coefficients):
        self._extrapolation = [float(c) for c in coefficients]

    def getCutoffDistance(self):
        return self._cutoff

    def setCutoffDistance(self, distance):
        self._cutoff = float(distance)

    def getPMEParameters(self):
        return (self._alpha, self._grid[0], self._grid[1], self._grid[2])

    def setPMEParameters(self, alpha, nx, ny, nz):
        self._alpha = float(alpha)
        self._grid = [int(nx), int(ny), int(nz)]

    def getAEwald(self):
        return self._alpha

    def setAEwald(self, alpha):
        self._alpha = float(alpha)

    def getPmeBSplineOrder(self):
        # hard-wired in the reference (MPIDForce.cpp:43, MPIDReferenceForce.cpp:2543)
        return 6

    def getPmeGridDimensions(self):
        return list(self._grid)

    def setPmeGridDimensions(self, dims):
        self._grid = [int(d) for d in dims]

    def getMutualInducedMaxIterations(self):
        return self._max_iter

    def setMutualInducedMaxIterations(self, n):
        self._max_iter = int(n)

    def getMutualInducedTargetEpsilon(self):
        return self._target_eps

    def setMutualInducedTargetEpsilon(self, eps):
        self._target_eps = float(eps)

    def getEwaldErrorTolerance(self):
        return self._ewald_tol

    def setEwaldErrorTolerance(self, tol):
        self._ewald_tol = float(tol)

    def get14ScaleFactor(self):
        return self._scale14

    def set14ScaleFactor(self, f):
        self._scale14 = float(f)

    def getDefaultTholeWidth(self):
        return self._default_thole

    def setDefaultTholeWidth(self, w):
        self._default_thole = float(w)

    def getForceGroup(self):
        return self._force_group

    def setForceGroup(self, g):
        self._force_group = int(g)

    def usesPeriodicBoundaryConditions(self):
        return self._nonbonded_method == MPIDForce.PME

    # ---- per-particle parameters ----------------------------------------------
    def addMultipole(self, charge, molecularDipole, molecularQuadrupole, molecularOctopole,
                     axisType, multipoleAtomZ, multipoleAtomX, multipoleAtomY, thole, alphas):
        rec = self._record(charge, molecularDipole, molecularQuadrupole, molecularOctopole,
                           axisType, multipoleAtomZ, multipoleAtomX, multipoleAtomY, thole, alphas)
        self._multipoles.append(rec)
        self._covalent.append([[] for _ in range(MPIDForce.CovalentEnd)])
        return len(self._multipoles) - 1

    @staticmethod
    def _record(charge, dip, quad, octo, axisType, z, x, y, thole, alphas):
        dip = [float(v) for v in dip]
        quad = [float(v) for v in quad]
        octo = [float(v) for v in octo]
        alphas = [float(v) for v in alphas]
        if len(dip) != 3 or len(quad) != 6 or len(octo) != 10 or len(alphas) != 3:
            raise MPIDForceError("MPIDForce: wrong number of multipole or polarizability components")
        return dict(charge=float(charge), dipole=dip, quadrupole=quad, octopole=octo,
                    axisType=int(axisType), atomZ=int(z), atomX=int(x), atomY=int(y),
                    thole=float(thole), alphas=alphas)

    def getMultipoleParameters(self, index):
        m = self._multipoles[index]
        return (m["charge"], list(m["dipole"]), list(m["quadrupole"]), list(m["octopole"]),
                m["axisType"], m["atomZ"], m["atomX"], m["atomY"], m["thole"], list(m["alphas"]))

    def setMultipoleParameters(self, index, charge, molecularDipole, molecularQuadrupole, molecularOctopole,
                               axisType, multipoleAtomZ, multipoleAtomX, multipoleAtomY, thole, alphas):
        self._multipoles[index] = self._record(charge, molecularDipole, molecularQuadrupole, molecularOctopole,
                                               axisType, multipoleAtomZ, multipoleAtomX, multipoleAtomY,
                                               thole, alphas)

    def setCovalentMap(self, index, typeId, covalentAtoms):
        self._covalent[index][int(typeId)] = [int(a) for a in covalentAtoms]

    def getCovalentMap(self, index, typeId):
        return list(self._covalent[index][int(typeId)])

    def getCovalentMaps(self, index):
        return [list(lst) for lst in self._covalent[index]]

    # ---- flattened arrays for the C ABI ----------------------------------------
    def flatten(self):
        """Flat numpy arrays in the layout of ``include/mpid_hip.h`` (and of the oracle).

        Mirrors ReferenceCalcMPIDForceKernel::initialize
        (platforms/reference/src/MPIDReferenceKernels.cpp:84-177).
        """
        n = len(self._multipoles)
        out = dict(
            n=n,
            charges=np.array([m["charge"] for m in self._multipoles], dtype=np.float64),
            dipoles=np.array([m["dipole"] for m in self._multipoles], dtype=np.float64).reshape(-1),
            quadrupoles=np.array([m["quadrupole"] for m in self._multipoles], dtype=np.float64).reshape(-1),
            octopoles=np.array([m["octopole"] for m in self._multipoles], dtype=np.float64).reshape(-1),
            axis_types=np.array([m["axisType"] for m in self._multipoles], dtype=np.int32),
            atom_z=np.array([m["atomZ"] for m in self._multipoles], dtype=np.int32),
            atom_x=np.array([m["atomX"] for m in self._multipoles], dtype=np.int32),
            atom_y=np.array([m["atomY"] for m in self._multipoles], dtype=np.int32),
            tholes=np.array([m["thole"] for m in self._multipoles], dtype=np.float64),
            polarity=np.array([m["alphas"] for m in self._multipoles], dtype=np.float64).reshape(-1),
        )
        offsets = [0]
        atoms = []
        for a in range(n):
            for t in range(4):  # only Covalent12..15 enter the scaling (MPIDReferenceForce.cpp:215)
                atoms.extend(self._covalent[a][t])
                offsets.append(len(atoms))
        out["cov_offsets"] = np.array(offsets, dtype=np.int32)
        out["cov_atoms"] = np.array(atoms if atoms else [0], dtype=np.int32)
        out.update(
            nonbonded_method=self._nonbonded_method,
            polarization=self._polarization_type,
            cutoff=self._cutoff,
            alpha=self._alpha,
            grid=tuple(self._grid),
            default_thole=self._default_thole,
            scale14=self._scale14,
            max_iter=self._max_iter,
            target_eps=self._target_eps,
            extrapolation_coefficients=np.array(self._extrapolation, dtype=np.float64),
        )
        return out
