"""The C ABI library loads and exports every symbol include/ivp.h declares (no GPU calls)."""

import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "ivp.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(ivp_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_entry_points():
    names = _declared()
    assert "ivp_ip_solve" in names and "ivp_neighbors_build" in names and "ivp_step" in names


def test_library_exports_every_declared_symbol():
    from implicitvelocityprojection_amd import _native as N
    lib = N.load()
    for name in _declared():
        assert hasattr(lib, name), name
    assert set(_declared()) == set(N.SIGNATURES), "ctypes signatures must cover exactly the header"


def test_status_strings_without_gpu():
    from implicitvelocityprojection_amd import _native as N
    lib = N.load()
    assert lib.ivp_version().decode().startswith("ivp")
    for s in (0, 1, 2, 3, -1, -2, -3, -4, -5, -6):
        assert lib.ivp_status_string(s)
