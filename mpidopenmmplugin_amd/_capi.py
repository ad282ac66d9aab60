"""ctypes binding of ``libmpid_hip.so`` (declarations in include/mpid_hip.h).

The library is the product: there is no CPU fallback.  Importing the binding
fails loudly when the library is missing; calls fail when no GPU is present.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# MPID_PME_ORDER=5 in the environment selects the library built for B-spline
# order 5 (a compile-time option; the reference and the default build use 6)
PME_ORDER = 5 if os.environ.get("MPID_PME_ORDER", "6") == "5" else 6
LIB_PATH = os.path.join(_HERE, "libmpid_hip_order5.so" if PME_ORDER == 5 else "libmpid_hip.so")

MPID_OK = 0
MPID_ERR_INVALID = 1
MPID_ERR_NOT_CONVERGED = 2
MPID_ERR_HIP = 3
MPID_ERR_UNSUPPORTED = 4

MPID_NOALLOC_MAX_POINTS = 65536

MPID_PRECISION_MIXED = 0
MPID_PRECISION_DOUBLE = 1

PROFILE_NAMES = ["sort", "rotate", "tiles", "spread", "fft", "convolve", "gather", "fixed_pairs",
                 "induced_pairs", "electrostatic_pairs", "recip_atoms", "solver", "other", "covalent_pairs",
                 "refinement"]

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


class Options(ctypes.Structure):
    _fields_ = [
        ("nonbonded_method", ctypes.c_int),
        ("polarization", ctypes.c_int),
        ("cutoff", ctypes.c_double),
        ("ewald_alpha", ctypes.c_double),
        ("grid", ctypes.c_int * 3),
        ("default_thole", ctypes.c_double),
        ("scale14", ctypes.c_double),
        ("max_iterations", ctypes.c_int),
        ("target_epsilon", ctypes.c_double),
        ("precision", ctypes.c_int),
        ("flags", ctypes.c_int),
        ("n_extrapolation", ctypes.c_int),
        ("extrapolation_coefficients", _dp),
        ("spread_level", ctypes.c_int),
        ("spread_mode", ctypes.c_int),
    ]


MPID_FLAG_CONSERVATIVE_TORQUES = 1
MPID_FLAG_SINGLE_STREAM = 2
MPID_FLAG_DIIS = 4


class Parameters(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int),
        ("charges", _dp), ("dipoles", _dp), ("quadrupoles", _dp), ("octopoles", _dp),
        ("axis_types", _ip), ("atom_z", _ip), ("atom_x", _ip), ("atom_y", _ip),
        ("tholes", _dp), ("polarity", _dp),
        ("covalent_offsets", _ip), ("covalent_atoms", _ip),
    ]


class Stats(ctypes.Structure):
    _fields_ = [
        ("iterations", ctypes.c_int),
        ("epsilon", ctypes.c_double),
        ("converged", ctypes.c_int),
        ("refinements", ctypes.c_int),
        ("matvecs", ctypes.c_int),
        ("ntiles", ctypes.c_int),
        ("energy_terms", ctypes.c_double * 3),
    ]


class Gradients(ctypes.Structure):
    _fields_ = [
        ("positions", _dp), ("charges", _dp), ("dipoles", _dp), ("quadrupoles", _dp), ("octopoles", _dp),
        ("polarity", _dp), ("tholes", _dp), ("default_thole", _dp),
        ("adjoint_iterations", ctypes.c_int),
    ]


MPID_GRADIENT_ALL_SITES = 1

# name -> (restype, argtypes); the full exported surface of include/mpid_hip.h
SIGNATURES = {
    "mpid_create": (ctypes.c_int, [ctypes.POINTER(Parameters), ctypes.POINTER(Options),
                                   ctypes.POINTER(ctypes.c_void_p)]),
    "mpid_update_parameters": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(Parameters)]),
    "mpid_execute": (ctypes.c_int, [ctypes.c_void_p, _dp, _dp, ctypes.c_int, ctypes.c_int, _dp, _dp]),
    "mpid_execute_device": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, _dp, ctypes.c_void_p, _dp]),
    "mpid_get_induced_dipoles": (ctypes.c_int, [ctypes.c_void_p, _dp, _dp, _dp]),
    "mpid_get_electrostatic_potential": (ctypes.c_int, [ctypes.c_void_p, _dp, _dp, ctypes.c_int, _dp, _dp]),
    "mpid_get_electrostatic_potential_noalloc": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _dp, _dp]),
    "mpid_get_system_multipole_moments": (ctypes.c_int, [ctypes.c_void_p, _dp, _dp, _dp, _dp]),
    "mpid_get_lab_frame_permanent_dipoles": (ctypes.c_int, [ctypes.c_void_p, _dp, _dp, _dp]),
    "mpid_get_total_dipoles": (ctypes.c_int, [ctypes.c_void_p, _dp, _dp, _dp]),
    "mpid_get_pme_parameters": (ctypes.c_int, [ctypes.c_void_p, _dp, _ip, _ip, _ip]),
    "mpid_spread_tilings": (ctypes.c_int, [_ip, ctypes.c_int, _ip, _ip]),
    "mpid_get_pme_bspline_order": (ctypes.c_int, []),
    "mpid_compute_fixed_field": (ctypes.c_int, [ctypes.c_void_p, _dp, _dp, _dp]),
    "mpid_compute_induced_field": (ctypes.c_int, [ctypes.c_void_p, _dp, _dp, _dp, _dp]),
    "mpid_energy_gradients": (ctypes.c_int, [ctypes.c_void_p, _dp, _dp, ctypes.c_int, _dp,
                                             ctypes.POINTER(Gradients)]),
    "mpid_get_stats": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(Stats)]),
    "mpid_set_profiling": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "mpid_set_profile_mask": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint]),
    "mpid_get_profile": (ctypes.c_int, [ctypes.c_void_p, _dp, _ip, ctypes.c_int]),
    "mpid_last_error": (ctypes.c_char_p, [ctypes.c_void_p]),
    "mpid_create_error": (ctypes.c_char_p, []),
    "mpid_destroy": (None, [ctypes.c_void_p]),
}

_lib = None


def load():
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                              "(make -C mpidopenmmplugin_amd/csrc)")
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def dptr(a):
    return a.ctypes.data_as(_dp)


def iptr(a):
    return a.ctypes.data_as(_ip)


def make_parameters(flat):
    """Parameters struct from ``MPIDForce.flatten()``; returns (struct, keep-alive dict)."""
    keep = {}
    for name, dt in (("charges", np.float64), ("dipoles", np.float64), ("quadrupoles", np.float64),
                     ("octopoles", np.float64), ("axis_types", np.int32), ("atom_z", np.int32),
                     ("atom_x", np.int32), ("atom_y", np.int32), ("tholes", np.float64),
                     ("polarity", np.float64), ("cov_offsets", np.int32), ("cov_atoms", np.int32)):
        keep[name] = np.ascontiguousarray(flat[name], dtype=dt)
    p = Parameters()
    p.n = int(flat["n"])
    p.charges = dptr(keep["charges"])
    p.dipoles = dptr(keep["dipoles"])
    p.quadrupoles = dptr(keep["quadrupoles"])
    p.octopoles = dptr(keep["octopoles"])
    p.axis_types = iptr(keep["axis_types"])
    p.atom_z = iptr(keep["atom_z"])
    p.atom_x = iptr(keep["atom_x"])
    p.atom_y = iptr(keep["atom_y"])
    p.tholes = dptr(keep["tholes"])
    p.polarity = dptr(keep["polarity"])
    p.covalent_offsets = iptr(keep["cov_offsets"])
    p.covalent_atoms = iptr(keep["cov_atoms"])
    return p, keep


def make_options(flat, precision, flags=0, spread_tiling=(0, 0)):
    """spread_tiling: (spread_level, spread_mode), (0, 0) for the library's choice."""
    o = Options()
    o.nonbonded_method = int(flat["nonbonded_method"])
    o.polarization = int(flat["polarization"])
    o.cutoff = float(flat["cutoff"])
    o.ewald_alpha = float(flat["alpha"])
    for k in range(3):
        o.grid[k] = int(flat["grid"][k])
    o.default_thole = float(flat["default_thole"])
    o.scale14 = float(flat["scale14"])
    o.max_iterations = int(flat["max_iter"])
    o.target_epsilon = float(flat["target_eps"])
    o.precision = int(precision)
    o.flags = int(flags)
    coeffs = flat.get("extrapolation_coefficients")
    if coeffs is not None and len(coeffs) > 0:
        c = np.ascontiguousarray(coeffs, dtype=np.float64)
        o._keep = c  # the library copies the array during the call that takes the options
        o.n_extrapolation = int(c.size)
        o.extrapolation_coefficients = c.ctypes.data_as(_dp)
    o.spread_level = int(spread_tiling[0])
    o.spread_mode = int(spread_tiling[1])
    return o


def spread_tilings(grid):
    """The (spread_level, spread_mode) pairs the library offers for a PME grid,
    its built-in choice first (mpid_spread_tilings; no GPU needed)."""
    lib = load()
    g = (ctypes.c_int * 3)(*[int(x) for x in grid])
    lv = (ctypes.c_int * 32)()
    md = (ctypes.c_int * 32)()
    count = lib.mpid_spread_tilings(g, 32, lv, md)
    if count < 0:
        raise ValueError(f"no spreading tilings for PME grid {tuple(grid)} (error {count})")
    return [(int(lv[i]), int(md[i])) for i in range(min(count, 32))]
