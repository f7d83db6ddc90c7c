"""The C-ABI library loads (no GPU needed) and exports every function that
include/pdp3_hip.h declares; the Python binding covers the same set."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "pdp3_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pdp3_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_functions():
    names = declared_functions()
    assert "pdp3_step" in names and "pdp3_ctx_create" in names
    assert len(names) >= 30


def test_library_exports_every_declared_symbol():
    from pdp3_amd import capi
    lib = ctypes.CDLL(capi.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_matches_header():
    from pdp3_amd import capi
    assert sorted(capi.SIGNATURES) == declared_functions()


def test_version_string():
    from pdp3_amd import capi
    assert b"gfx950" in capi.lib().pdp3_version()
