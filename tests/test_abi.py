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


def _oracle_imports(path):
    """(line, module) of every import of the oracle package in a Python file."""
    import ast
    with open(path) as fh:
        tree = ast.parse(fh.read(), path)
    out = []
    for node in ast.walk(tree):
        if isinstance(node, ast.Import):
            out += [(node.lineno, a.name) for a in node.names if a.name.split(".")[0] == "oracle"]
        elif isinstance(node, ast.ImportFrom) and node.module and node.module.split(".")[0] == "oracle":
            out.append((node.lineno, node.module))
    return out


def test_product_never_imports_the_oracle():
    """The package is the product and never reaches the FP64 checker; bench.py imports it only
    in its CPU-baseline leg, after the timed GPU requests."""
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    pkg = os.path.join(root, "dynlp_amd")
    for d, _, files in os.walk(pkg):
        for f in files:
            if f.endswith(".py"):
                assert not _oracle_imports(os.path.join(d, f)), os.path.join(d, f)
    bench = os.path.join(root, "bench.py")
    with open(bench) as fh:
        lines = fh.read().splitlines()
    marker = next(i for i, s in enumerate(lines, 1) if "CPU baseline" in s and s.strip().startswith("#"))
    imports = _oracle_imports(bench)
    assert imports and all(line > marker for line, _ in imports), (marker, imports)
