"""The C ABI library is built, loads, and exports every entry point that
include/mpid_hip.h declares (no compute calls: these run without a GPU)."""
import ctypes
import os
import re

from mpidopenmmplugin_amd import _capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "mpid_hip.h")).read()
    return sorted(set(re.findall(r"\b(mpid_[a-z_]+)\s*\(", text)))


def test_header_declares_the_boundary():
    syms = declared_symbols()
    for required in ("mpid_create", "mpid_execute", "mpid_execute_device", "mpid_get_induced_dipoles",
                     "mpid_update_parameters", "mpid_get_pme_parameters", "mpid_destroy"):
        assert required in syms


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_capi.LIB_PATH)
    for name in declared_symbols():
        assert hasattr(lib, name), name
    # and the binding covers all of them
    assert set(declared_symbols()) == set(_capi.SIGNATURES)


def test_binding_loads():
    lib = _capi.load()
    assert lib.mpid_create_error() is not None
