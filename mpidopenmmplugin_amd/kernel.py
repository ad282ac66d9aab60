"""The gfx950 counterpart of ``CalcMPIDForceKernel`` and a small context object.

``HipCalcMPIDForceKernel`` mirrors the reference's platform-kernel interface
(openmmapi/include/openmm/mpidKernels.h:50-104): ``initialize``, ``execute``,
the dipole getters, ``copyParametersToContext`` and ``getPMEParameters``.
``MPIDContext`` stands in for the OpenMM ``Context`` that drives a kernel in
the reference's tests (TestReferenceMPIDForce.cpp: ``context.setPositions``,
``context.getState(Energy|Forces)``).  All arithmetic happens in
``libmpid_hip.so``; nothing here computes physics.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _capi, autotune
from .mpidforce import MPIDForce, MPIDForceError, automatic_pme_parameters


class MPIDConvergenceError(MPIDForceError):
    """Mutual induced dipoles did not converge (MPIDReferenceForce.cpp:2229-2235)."""


class MPIDUnsupportedError(MPIDForceError):
    pass


_PRECISIONS = {"mixed": _capi.MPID_PRECISION_MIXED, "double": _capi.MPID_PRECISION_DOUBLE}


def _box9(box):
    if box is None:  # NoCutoff takes no box; PME rejects a zero box
        return np.zeros(9)
    b = np.asarray(box, dtype=np.float64)
    if b.size == 3:
        b = np.diag(b.reshape(3))
    return np.ascontiguousarray(b.reshape(9))


class HipCalcMPIDForceKernel:
    """gfx950 implementation of CalcMPIDForceKernel behind a C ABI."""

    Name = "CalcMPIDForce"

    def __init__(self, precision="mixed", conservative_torques=False, single_stream=False, spread_tiling=None,
                 solver="cg"):
        """``conservative_torques`` restores the half of the induced-induced torque
        about the pair axis that the reference omits
        (MPIDReferenceForce.cpp:4854-4855); the default reproduces the reference.
        ``single_stream`` issues all work on one HIP stream (for many small
        contexts evaluated at once).  ``spread_tiling``: None takes the PME
        spreading's tiling from the autotuner's table (the library's own
        choice for grids not in it), "library" the library's choice, or a
        (spread_level, spread_mode) pair of ``_capi.spread_tilings(grid)``.
        ``solver``: "cg" (preconditioned conjugate gradients) or "diis" (the
        reference's DIIS iteration, for parity of the iterates; double
        precision only)."""
        if precision not in _PRECISIONS:
            raise ValueError(f"precision must be one of {sorted(_PRECISIONS)}")
        self.precision = precision
        self.spread_tiling = spread_tiling
        self.flags = _capi.MPID_FLAG_CONSERVATIVE_TORQUES if conservative_torques else 0
        if single_stream:
            self.flags |= _capi.MPID_FLAG_SINGLE_STREAM
        if solver not in ("cg", "diis"):
            raise ValueError("solver must be 'cg' or 'diis'")
        if solver == "diis":
            self.flags |= _capi.MPID_FLAG_DIIS
        self._lib = _capi.load()
        self._ctx = ctypes.c_void_p()
        self._n = 0

    # -- error handling ----------------------------------------------------------
    def _check(self, rc):
        if rc == _capi.MPID_OK:
            return
        msg = self._lib.mpid_last_error(self._ctx).decode()
        if rc == _capi.MPID_ERR_NOT_CONVERGED:
            raise MPIDConvergenceError(msg)
        if rc == _capi.MPID_ERR_UNSUPPORTED:
            raise MPIDUnsupportedError(msg)
        if rc == _capi.MPID_ERR_INVALID:
            raise MPIDForceError(msg)
        raise RuntimeError(f"libmpid_hip error {rc}: {msg}")

    # -- CalcMPIDForceKernel ------------------------------------------------------
    def initialize(self, system_num_particles, force: MPIDForce, default_box=None):
        """default_box: the System's default periodic box (edges or vectors,
        nm); needed only by PME with alpha or the grid left at 0, which
        takes both from the Ewald error tolerance
        (MPIDCudaKernels.cpp:441-448)."""
        if force.getNumMultipoles() != system_num_particles:
            raise MPIDForceError("MPIDForce must have exactly as many particles as the System it belongs to.")
        flat = force.flatten()
        if flat["nonbonded_method"] == MPIDForce.PME and (flat["alpha"] == 0.0 or flat["grid"][0] == 0):
            if default_box is None:
                raise MPIDForceError("PME with alpha or grid 0 takes them from the Ewald error tolerance and "
                                     "needs the default box")
            a, nx, ny, nz = automatic_pme_parameters(flat["cutoff"], force.getEwaldErrorTolerance(), default_box)
            flat["alpha"], flat["grid"] = a, (nx, ny, nz)
        params, keep = _capi.make_parameters(flat)
        tiling = (0, 0)
        if flat["nonbonded_method"] == MPIDForce.PME:
            if self.spread_tiling is None:
                tiling = autotune.lookup(flat["grid"]) or (0, 0)
            elif self.spread_tiling != "library":
                tiling = tuple(int(x) for x in self.spread_tiling)
        opts = _capi.make_options(flat, _PRECISIONS[self.precision], self.flags, tiling)
        ctx = ctypes.c_void_p()
        rc = self._lib.mpid_create(ctypes.byref(params), ctypes.byref(opts), ctypes.byref(ctx))
        if rc != _capi.MPID_OK:
            msg = self._lib.mpid_create_error().decode()
            if rc == _capi.MPID_ERR_UNSUPPORTED:
                raise MPIDUnsupportedError(msg)
            if rc == _capi.MPID_ERR_INVALID:
                raise MPIDForceError(msg)
            raise RuntimeError(f"libmpid_hip error {rc}: {msg}")
        self._ctx = ctx
        self._n = flat["n"]
        self._pme = (flat["alpha"], *flat["grid"])

    def execute(self, positions, box, includeForces=True, includeEnergy=True, forces=None):
        """Energy (kJ/mol) and forces (kJ/mol/nm, [n,3]); forces are added to ``forces`` if given."""
        pos = np.ascontiguousarray(positions, dtype=np.float64).reshape(-1)
        f = np.zeros(3 * self._n) if forces is None else np.ascontiguousarray(forces, dtype=np.float64).reshape(-1)
        e = ctypes.c_double(0.0)
        b = _box9(box)
        self._check(self._lib.mpid_execute(self._ctx, _capi.dptr(pos), _capi.dptr(b), int(includeForces),
                                           int(includeEnergy), _capi.dptr(f), ctypes.byref(e)))
        return e.value, f.reshape(self._n, 3)

    def execute_device(self, d_positions_ptr, box, d_forces_ptr=None):
        """Positions/forces already in device memory (raw HIP pointers)."""
        e = ctypes.c_double(0.0)
        b = _box9(box)
        self._check(self._lib.mpid_execute_device(self._ctx, ctypes.c_void_p(d_positions_ptr), _capi.dptr(b),
                                                  ctypes.c_void_p(d_forces_ptr or 0), ctypes.byref(e)))
        return e.value

    def _getter(self, fn, positions, box):
        pos = np.ascontiguousarray(positions, dtype=np.float64).reshape(-1)
        out = np.zeros(3 * self._n)
        b = _box9(box)
        self._check(fn(self._ctx, _capi.dptr(pos), _capi.dptr(b), _capi.dptr(out)))
        return out.reshape(self._n, 3)

    def getInducedDipoles(self, positions, box):
        return self._getter(self._lib.mpid_get_induced_dipoles, positions, box)

    def getLabFramePermanentDipoles(self, positions, box):
        return self._getter(self._lib.mpid_get_lab_frame_permanent_dipoles, positions, box)

    def getTotalDipoles(self, positions, box):
        return self._getter(self._lib.mpid_get_total_dipoles, positions, box)

    def getElectrostaticPotential(self, points, positions, box):
        """CalcMPIDForceKernel::getElectrostaticPotential (mpidKernels.h:83-84):
        kJ/mol/e at points [m,3] (nm) from charges, dipoles + induced dipoles and
        quadrupoles, after an evaluation at ``positions``."""
        pos = np.ascontiguousarray(positions, dtype=np.float64).reshape(-1)
        pts = np.ascontiguousarray(points, dtype=np.float64).reshape(-1)
        out = np.zeros(pts.size // 3)
        b = _box9(box)
        self._check(self._lib.mpid_get_electrostatic_potential(self._ctx, _capi.dptr(pos), _capi.dptr(b),
                                                                int(out.size), _capi.dptr(pts), _capi.dptr(out)))
        return out

    def getSystemMultipoleMoments(self, positions, box, masses):
        """CalcMPIDForceKernel::getSystemMultipoleMoments (mpidKernels.h:86): 13
        values (charge e, dipole Debye, traceless quadrupole Buckingham, 3x3)
        about the centre of mass of ``masses``."""
        pos = np.ascontiguousarray(positions, dtype=np.float64).reshape(-1)
        m = np.ascontiguousarray(masses, dtype=np.float64).reshape(-1)
        if m.size != self._n:
            raise MPIDForceError("one mass per particle is required")
        out = np.zeros(13)
        b = _box9(box)
        self._check(self._lib.mpid_get_system_multipole_moments(self._ctx, _capi.dptr(pos), _capi.dptr(b),
                                                                 _capi.dptr(m), _capi.dptr(out)))
        return out

    def copyParametersToContext(self, force: MPIDForce):
        flat = force.flatten()
        params, keep = _capi.make_parameters(flat)
        self._check(self._lib.mpid_update_parameters(self._ctx, ctypes.byref(params)))

    def updateParameterArrays(self, flat):
        """copyParametersToContext from flat arrays (the layout of
        ``MPIDForce.flatten()``): charges, dipoles, quadrupoles, octopoles,
        polarity and tholes may change, the frames and covalent maps not."""
        params, keep = _capi.make_parameters(flat)
        self._check(self._lib.mpid_update_parameters(self._ctx, ctypes.byref(params)))

    def energyGradients(self, positions, box, all_sites=False):
        """Energy (kJ/mol) and its gradients with respect to the positions and
        the per-particle parameters (mpid_energy_gradients): a dict of numpy
        arrays in the layouts of ``MPIDForce.flatten()`` -- positions [n,3],
        charges [n], dipoles [n,3], quadrupoles [n,6], octopoles [n,10],
        polarity [n,3], tholes [n] -- plus default_thole (float), energy and
        adjoint_iterations.  ``all_sites``: also the gradients of the higher
        moments and polarisabilities of bare-charge sites."""
        n = self._n
        pos = np.ascontiguousarray(positions, dtype=np.float64).reshape(-1)
        out = dict(positions=np.zeros((n, 3)), charges=np.zeros(n), dipoles=np.zeros((n, 3)),
                   quadrupoles=np.zeros((n, 6)), octopoles=np.zeros((n, 10)), polarity=np.zeros((n, 3)),
                   tholes=np.zeros(n))
        dth = np.zeros(1)
        g = _capi.Gradients()
        for name, arr in out.items():
            setattr(g, name, _capi.dptr(arr))
        g.default_thole = _capi.dptr(dth)
        e = ctypes.c_double(0.0)
        b = _box9(box)
        flags = _capi.MPID_GRADIENT_ALL_SITES if all_sites else 0
        self._check(self._lib.mpid_energy_gradients(self._ctx, _capi.dptr(pos), _capi.dptr(b), flags,
                                                    ctypes.byref(e), ctypes.byref(g)))
        out["default_thole"] = float(dth[0])
        out["energy"] = e.value
        out["adjoint_iterations"] = int(g.adjoint_iterations)
        return out

    def getPMEParameters(self):
        a = ctypes.c_double()
        nx, ny, nz = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        self._check(self._lib.mpid_get_pme_parameters(self._ctx, ctypes.byref(a), ctypes.byref(nx),
                                                      ctypes.byref(ny), ctypes.byref(nz)))
        return a.value, nx.value, ny.value, nz.value

    # -- inspection ----------------------------------------------------------------
    def computeFixedField(self, positions, box):
        return self._getter(self._lib.mpid_compute_fixed_field, positions, box)

    def computeInducedField(self, positions, box, mu):
        pos = np.ascontiguousarray(positions, dtype=np.float64).reshape(-1)
        m = np.ascontiguousarray(mu, dtype=np.float64).reshape(-1)
        out = np.zeros(3 * self._n)
        b = _box9(box)
        self._check(self._lib.mpid_compute_induced_field(self._ctx, _capi.dptr(pos), _capi.dptr(b), _capi.dptr(m),
                                                         _capi.dptr(out)))
        return out.reshape(self._n, 3)

    def stats(self):
        s = _capi.Stats()
        self._check(self._lib.mpid_get_stats(self._ctx, ctypes.byref(s)))
        return dict(iterations=s.iterations, epsilon=s.epsilon, converged=bool(s.converged),
                    refinements=s.refinements, matvecs=s.matvecs, ntiles=s.ntiles,
                    energy_terms=list(s.energy_terms))

    def setProfiling(self, enable=True):
        self._check(self._lib.mpid_set_profiling(self._ctx, int(enable)))

    def setProfileFamilies(self, names):
        """Time only the named kernel families (see _capi.PROFILE_NAMES);
        unless one of them runs inside a solver iteration, the iterations
        keep their graph replays."""
        mask = 0
        for name in names:
            mask |= 1 << _capi.PROFILE_NAMES.index(name)
        self._check(self._lib.mpid_set_profile_mask(self._ctx, mask))

    def profile(self, reset=False):
        t = np.zeros(len(_capi.PROFILE_NAMES))
        c = np.zeros(len(_capi.PROFILE_NAMES), dtype=np.int32)
        self._check(self._lib.mpid_get_profile(self._ctx, _capi.dptr(t), _capi.iptr(c), int(reset)))
        return {name: (float(t[i]), int(c[i])) for i, name in enumerate(_capi.PROFILE_NAMES)}

    def __del__(self):
        try:
            if self._ctx:
                self._lib.mpid_destroy(self._ctx)
                self._ctx = ctypes.c_void_p()
        except Exception:
            pass


class State:
    def __init__(self, energy=None, forces=None):
        self._energy = energy
        self._forces = forces

    def getPotentialEnergy(self):
        return self._energy

    def getForces(self, asNumpy=True):
        return self._forces


class MPIDContext:
    """Minimal stand-in for an OpenMM Context holding one MPIDForce on the gfx950 kernel."""

    def __init__(self, force: MPIDForce, box, precision="mixed", conservative_torques=False):
        self.force = force
        self.box = _box9(box)
        self.kernel = HipCalcMPIDForceKernel(precision, conservative_torques)
        self.kernel.initialize(force.getNumMultipoles(), force, default_box=box)
        self.positions = None

    def setPositions(self, positions):
        self.positions = np.ascontiguousarray(positions, dtype=np.float64).reshape(-1, 3)

    def setPeriodicBoxVectors(self, a, b, c):
        self.box = _box9(np.array([a, b, c], dtype=np.float64))

    def getState(self, getEnergy=True, getForces=True):
        e, f = self.kernel.execute(self.positions, self.box, getForces, getEnergy)
        return State(e if getEnergy else None, f if getForces else None)

    def getInducedDipoles(self):
        return self.kernel.getInducedDipoles(self.positions, self.box)

    def getLabFramePermanentDipoles(self):
        return self.kernel.getLabFramePermanentDipoles(self.positions, self.box)

    def getTotalDipoles(self):
        return self.kernel.getTotalDipoles(self.positions, self.box)

    def getElectrostaticPotential(self, points):
        """MPIDForce::getElectrostaticPotential (MPIDForce.h:404)."""
        return self.kernel.getElectrostaticPotential(points, self.positions, self.box)

    def getSystemMultipoleMoments(self, masses):
        """MPIDForce::getSystemMultipoleMoments (MPIDForce.h:423); the masses are
        the System's particle masses."""
        return self.kernel.getSystemMultipoleMoments(self.positions, self.box, masses)

    def updateParametersInContext(self):
        self.kernel.copyParametersToContext(self.force)

    def getPMEParametersInContext(self):
        """MPIDForce::getPMEParametersInContext (MPIDForce.cpp:126-128):
        (alpha, nx, ny, nz) in use, also when taken from the tolerance."""
        return self.kernel.getPMEParameters()
