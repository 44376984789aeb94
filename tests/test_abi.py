"""The C-ABI library loads and exports every entry point include/lucent_hip.h declares."""
import re
from pathlib import Path

import pytest

from lucent_amd import _native

HEADER = Path(__file__).resolve().parents[1] / "include" / "lucent_hip.h"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lucent_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared_functions()
    assert "lucent_intersect" in names and "lucent_build_bvh" in names


def test_library_exports_every_declared_symbol():
    lib = _native.lib()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, f"not exported: {missing}"


def test_python_signatures_cover_header():
    assert set(declared_functions()) <= set(_native.SIGNATURES), \
        sorted(set(declared_functions()) - set(_native.SIGNATURES))


def test_version_and_error_plumbing():
    lib = _native.lib()
    assert b"gfx950" in lib.lucent_version()
    # A host-only call with a bad argument reports through lucent_last_error.
    import ctypes as C
    st = lib.lucent_build_bvh(None, 0, 4, 1, None, None, None, None)
    assert st == -1
    assert b"bins" in lib.lucent_last_error(None)
