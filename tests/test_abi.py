"""The C ABI library loads (no GPU needed) and exports every function include/dynlp.h declares."""
import ctypes as ct
import os
import re

from dynlp_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "dynlp.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(dlp_[a-z0-9_]+)\s*\(", src)))


def test_library_loads_and_exports_all_declared_symbols():
    lib = _lib.load()
    names = declared_functions()
    assert len(names) >= 10
    missing = [nm for nm in names if not hasattr(lib, nm)]
    assert not missing, missing


def test_binding_covers_declarations():
    assert set(declared_functions()) <= set(_lib.exported_symbols())


def test_abi_version_and_error_path():
    lib = _lib.load()
    assert lib.dlp_abi_version() >= 1
    # argument errors are reported without touching a device
    rc = lib.dlp_spmm_axpy(10, 0, None, None, None, None, 0, None, ct.c_double(1.0), None, 0, 0,
                           ct.c_double(0.0), None)
    assert rc == -1
    assert b"C=0" in lib.dlp_last_error()


def test_workspace_queries():
    lib = _lib.load()
    assert lib.dlp_cc_workspace_bytes(1000) >= 8000
    assert lib.dlp_scan_workspace_bytes(1 << 20) >= 4 << 20
