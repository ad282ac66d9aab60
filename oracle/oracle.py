"""ctypes front end of the FP64 CPU oracle (oracle/mpid_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product path never imports
this module.

The input is the flat parameter dictionary produced by
``MPIDForce.flatten()`` (same layout as the product's C ABI).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# MPID_PME_ORDER=5 in the environment selects the restatement built for
# B-spline order 5 (the reference's order is 6)
PME_ORDER = 5 if os.environ.get("MPID_PME_ORDER", "6") == "5" else 6
_LIB_PATH = os.path.join(_HERE, "libmpid_oracle_order5.so" if PME_ORDER == 5 else "libmpid_oracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


class _System(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int),
        ("charges", _dp), ("dipoles", _dp), ("quadrupoles", _dp), ("octopoles", _dp),
        ("axis_types", _ip), ("atom_z", _ip), ("atom_x", _ip), ("atom_y", _ip),
        ("tholes", _dp), ("polarity", _dp),
        ("cov_offsets", _ip), ("cov_atoms", _ip),
        ("polarization", ctypes.c_int),
        ("cutoff", ctypes.c_double), ("alpha", ctypes.c_double),
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
        ("default_thole", ctypes.c_double), ("scale14", ctypes.c_double),
        ("max_iter", ctypes.c_int), ("target_eps", ctypes.c_double),
        ("box", ctypes.c_double * 9),
        ("flags", ctypes.c_int),
        ("n_coeff", ctypes.c_int), ("coeffs", _dp),
        ("method", ctypes.c_int),
    ]


class _Out(ctypes.Structure):
    _fields_ = [
        ("lab_dipoles", _dp), ("lab_quads", _dp), ("lab_octs", _dp), ("lab_pol", _dp),
        ("fixed_field", _dp), ("induced", _dp), ("phi", _dp), ("phidp", _dp),
        ("torques", _dp), ("forces", _dp),
        ("energy_terms", ctypes.c_double * 4), ("energy", ctypes.c_double), ("energy_variational", ctypes.c_double),
        ("iterations", ctypes.c_int), ("eps", ctypes.c_double), ("converged", ctypes.c_int),
        ("error", ctypes.c_char * 256),
    ]


class OracleError(RuntimeError):
    pass


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.oracle_evaluate.argtypes = [ctypes.POINTER(_System), _dp, ctypes.POINTER(_Out)]
        L.oracle_evaluate.restype = ctypes.c_int
        L.oracle_induced_field.argtypes = [ctypes.POINTER(_System), _dp, _dp, _dp]
        L.oracle_induced_field.restype = ctypes.c_int
        L.oracle_potential.argtypes = [ctypes.POINTER(_System), _dp, ctypes.c_int, _dp, _dp]
        L.oracle_potential.restype = ctypes.c_int
        L.oracle_multipole_moments.argtypes = [ctypes.POINTER(_System), _dp, _dp, _dp]
        L.oracle_multipole_moments.restype = ctypes.c_int
        L.oracle_bspline.argtypes = [ctypes.c_double, _dp]
        L.oracle_bspline_moduli.argtypes = [ctypes.c_int, _dp]
        L.oracle_fft3d.argtypes = [_dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.oracle_torque_to_force.argtypes = [ctypes.POINTER(_System), _dp, _dp, _dp]
        L.oracle_torque_to_force.restype = ctypes.c_int
        _lib = L
    return _lib


def _ptr(a, kind):
    return a.ctypes.data_as(_dp if kind == "d" else _ip)


def _system(flat, box, polarization=None, flags=0):
    keep = {}

    def arr(name, dtype):
        a = np.ascontiguousarray(flat[name], dtype=dtype)
        keep[name] = a
        return a

    s = _System()
    s.n = int(flat["n"])
    s.charges = _ptr(arr("charges", np.float64), "d")
    s.dipoles = _ptr(arr("dipoles", np.float64), "d")
    s.quadrupoles = _ptr(arr("quadrupoles", np.float64), "d")
    s.octopoles = _ptr(arr("octopoles", np.float64), "d")
    s.axis_types = _ptr(arr("axis_types", np.int32), "i")
    s.atom_z = _ptr(arr("atom_z", np.int32), "i")
    s.atom_x = _ptr(arr("atom_x", np.int32), "i")
    s.atom_y = _ptr(arr("atom_y", np.int32), "i")
    s.tholes = _ptr(arr("tholes", np.float64), "d")
    s.polarity = _ptr(arr("polarity", np.float64), "d")
    s.cov_offsets = _ptr(arr("cov_offsets", np.int32), "i")
    s.cov_atoms = _ptr(arr("cov_atoms", np.int32), "i")
    pol = flat["polarization"] if polarization is None else polarization
    if pol not in (0, 1, 2):
        raise OracleError("unknown polarization type %r" % (pol,))
    s.polarization = int(pol)
    coeffs = np.ascontiguousarray(flat.get("extrapolation_coefficients", [-0.154, 0.017, 0.658, 0.474]),
                                  dtype=np.float64)
    keep["coeffs"] = coeffs
    s.n_coeff = int(coeffs.size)
    s.coeffs = _ptr(coeffs, "d")
    s.cutoff = float(flat["cutoff"])
    s.alpha = float(flat["alpha"])
    s.nx, s.ny, s.nz = (int(g) for g in flat["grid"])
    s.default_thole = float(flat["default_thole"])
    s.scale14 = float(flat["scale14"])
    s.max_iter = int(flat["max_iter"])
    s.target_eps = float(flat["target_eps"])
    b = np.zeros(9) if box is None else np.asarray(box, dtype=np.float64).reshape(-1)  # NoCutoff: no box
    if b.size == 3:
        b = np.diag(b).reshape(-1)
    for k in range(9):
        s.box[k] = float(b[k])
    s.flags = int(flags)
    s.method = int(flat.get("nonbonded_method", 1))
    return s, keep


def evaluate(flat, positions, box, polarization=None, details=False, conservative_torques=False):
    """Energy (kJ/mol), forces (kJ/mol/nm, [n,3]) and optional intermediates.

    ``conservative_torques`` adds the induced-induced torque term that the
    reference omits (MPIDReferenceForce.cpp:4854-4855)."""
    n = int(flat["n"])
    s, keep = _system(flat, box, polarization, 1 if conservative_torques else 0)
    pos = np.ascontiguousarray(positions, dtype=np.float64).reshape(-1)
    bufs = {}
    out = _Out()
    sizes = dict(lab_dipoles=3, lab_quads=6, lab_octs=10, lab_pol=6, fixed_field=3, induced=3,
                 phi=35, phidp=35, torques=3, forces=3)
    for name, k in sizes.items():
        if name == "forces" or details:
            bufs[name] = np.zeros(n * k, dtype=np.float64)
            setattr(out, name, _ptr(bufs[name], "d"))
    rc = lib().oracle_evaluate(ctypes.byref(s), _ptr(pos, "d"), ctypes.byref(out))
    if rc != 0:
        raise OracleError(out.error.decode())
    res = dict(energy=out.energy, energy_variational=out.energy_variational, forces=bufs["forces"].reshape(n, 3),
               energy_terms=list(out.energy_terms), iterations=out.iterations, eps=out.eps)
    if details:
        for name, k in sizes.items():
            res[name] = bufs[name].reshape(n, k)
    return res


def induced_field(flat, positions, box, mu):
    n = int(flat["n"])
    s, keep = _system(flat, box, 1)
    pos = np.ascontiguousarray(positions, dtype=np.float64).reshape(-1)
    m = np.ascontiguousarray(mu, dtype=np.float64).reshape(-1)
    out = np.zeros(3 * n, dtype=np.float64)
    rc = lib().oracle_induced_field(ctypes.byref(s), _ptr(pos, "d"), _ptr(m, "d"), _ptr(out, "d"))
    if rc != 0:
        raise OracleError("oracle_induced_field failed")
    return out.reshape(n, 3)


def bspline(w):
    th = np.zeros((PME_ORDER, 5), dtype=np.float64)
    lib().oracle_bspline(float(w), _ptr(th, "d"))
    return th


def bspline_moduli(n):
    m = np.zeros(n, dtype=np.float64)
    lib().oracle_bspline_moduli(int(n), _ptr(m, "d"))
    return m


def fft3d(a, sign=-1):
    a = np.ascontiguousarray(a, dtype=np.complex128)
    buf = np.ascontiguousarray(np.stack([a.real, a.imag], axis=-1), dtype=np.float64)
    lib().oracle_fft3d(_ptr(buf, "d"), a.shape[0], a.shape[1], a.shape[2], int(sign))
    return buf[..., 0] + 1j * buf[..., 1]


def electrostatic_potential(flat, positions, box, points):
    """getElectrostaticPotential (kJ/mol/e) at points [m,3] after the solve."""
    s, keep = _system(flat, box)
    pos = np.ascontiguousarray(positions, dtype=np.float64).reshape(-1)
    pts = np.ascontiguousarray(points, dtype=np.float64).reshape(-1)
    out = np.zeros(pts.size // 3)
    rc = lib().oracle_potential(ctypes.byref(s), _ptr(pos, "d"), int(out.size), _ptr(pts, "d"), _ptr(out, "d"))
    if rc != 0:
        raise OracleError("oracle_potential failed (%d)" % rc)
    return out


def system_multipole_moments(flat, positions, box, masses):
    """getSystemMultipoleMoments: 13 values (charge, Debye, Buckingham)."""
    s, keep = _system(flat, box)
    pos = np.ascontiguousarray(positions, dtype=np.float64).reshape(-1)
    m = np.ascontiguousarray(masses, dtype=np.float64).reshape(-1)
    out = np.zeros(13)
    rc = lib().oracle_multipole_moments(ctypes.byref(s), _ptr(pos, "d"), _ptr(m, "d"), _ptr(out, "d"))
    if rc != 0:
        raise OracleError("oracle_multipole_moments failed (%d)" % rc)
    return out


def torque_to_force(flat, positions, torques, conservative_torques=False):
    """mapTorqueToForce alone: the forces [n,3] that the frame chain rule
    gives for per-site torques [n,3] at the given positions
    (``conservative_torques``: ThreeFold frames by the exact chain rule)."""
    n = int(flat["n"])
    s, keep = _system(flat, None, flags=1 if conservative_torques else 0)
    pos = np.ascontiguousarray(positions, dtype=np.float64).reshape(-1)
    t = np.ascontiguousarray(torques, dtype=np.float64).reshape(-1)
    out = np.zeros(3 * n)
    rc = lib().oracle_torque_to_force(ctypes.byref(s), _ptr(pos, "d"), _ptr(t, "d"), _ptr(out, "d"))
    if rc != 0:
        raise OracleError("oracle_torque_to_force failed (%d)" % rc)
    return out.reshape(n, 3)
