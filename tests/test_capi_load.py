"""CPU-side checks of the C ABI library: it loads and exports every symbol
include/gp5.h declares.  No compute call is made (no GPU here)."""
import ctypes
import os
import re

from conftest import ROOT


def _declared():
    src = open(os.path.join(ROOT, "include", "gp5.h")).read()
    return sorted(set(re.findall(r"\b(gp5_[a-z_0-9]+)\s*\(", src)))


def test_header_and_binding_agree():
    from gp5_amd import capi

    assert sorted(capi.EXPORTS) == _declared()


def test_library_exports_every_declared_symbol():
    from gp5_amd import capi

    lib = ctypes.CDLL(capi.LIB_PATH)
    missing = [s for s in _declared() if not hasattr(lib, s)]
    assert not missing, missing


def test_library_reports_version_and_error_text():
    from gp5_amd import capi

    L = capi.load()
    assert L.gp5_version() >= 1
    # NULL context is rejected on the host without touching a device
    assert L.gp5_ctx_sync(None) == capi.GP5_ERR_ARG
    assert b"NULL" in L.gp5_last_error()
