"""The C-ABI library loads without a GPU, exports every entry point include/hotspice_hip.h
declares, and its struct layouts match the ctypes mirrors the host package uses."""
import os
import re

import pytest

from hotspice_amd import _lib

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "hotspice_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[a-z0-9_]+\s*\*?\s*(hs_[a-z0-9_]+)\s*\(", text, flags=re.M)))


def test_library_present():
    assert os.path.exists(_lib.LIB_PATH), "run __graft_entry__.build() first"


def test_exports_every_declared_symbol():
    lib = _lib.load()
    names = declared_functions()
    assert len(names) >= 12, names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) == set(_lib.EXPORTS)


def test_abi_layout_matches_ctypes():
    assert _lib.abi_layout_library() == _lib.abi_layout_python()


def test_version_and_error_string():
    lib = _lib.load()
    assert lib.hs_version() >= 100
    assert isinstance(lib.hs_last_error(), bytes)


def test_argument_errors_need_no_device():
    """Invalid arguments are rejected on the host before anything touches a device."""
    import ctypes as C
    lib = _lib.load()
    rc = lib.hs_philox_uniform53(1, 0, 1, -1, 10, 0, None, None)
    assert rc == -1 and b"philox" in lib.hs_last_error()
    inst = _lib.Instance()
    rc = lib.hs_recompute_fields(C.byref(inst), 1, None, 0, None)
    assert rc == -1 and b"grid" in lib.hs_last_error()
    g = _lib.KernelGeom()
    rc = lib.hs_build_dipolar_kernels(C.byref(g), None, None, None, None, None, None)
    assert rc == -1
