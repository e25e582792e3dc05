"""The C-ABI library loads and exports every function declared in
include/gfsdc.h (no GPU calls)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "gfsdc.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b([a-z_][a-z0-9_]*)\s*\(", src, re.M)
    return sorted(set(n for n in names if n.startswith(("gfsdc_", "shtns_"))))


def test_header_declares_both_tiers():
    names = declared_functions()
    assert len([n for n in names if n.startswith("shtns_")]) == 11
    assert "gfsdc_getdyntend" in names and "gfsdc_advance" in names


def test_library_exports_declared_symbols():
    import gfsdc
    lib = ctypes.CDLL(gfsdc.LIBPATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_loads_and_fills_defaults():
    import gfsdc
    c = gfsdc.cfg_defaults()
    assert c.nmem == 1 and c.massfix == 1 and c.physics == gfsdc.PHYS_HS
    assert c.rerth == float(__import__("numpy").float32(6.3712e6))
    assert gfsdc.load().gfsdc_last_error() is not None
