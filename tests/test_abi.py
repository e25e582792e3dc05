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


def test_cfg_layout_matches_bindings():
    """gfsdc_cfg has one layout: the library's sizeof, the ctypes Structure of
    the Python binding and the Fortran bind(c) type (gfsdc_dycore) agree."""
    import gfsdc
    lib = ctypes.CDLL(gfsdc.LIBPATH)
    lib.gfsdc_cfg_size.restype = ctypes.c_size_t
    assert lib.gfsdc_cfg_size() == ctypes.sizeof(gfsdc.Cfg) == 216
    drv = os.path.join(ROOT, "gfs-dycore_amd", "gfsdc_hs_driver")
    if os.path.exists(drv):   # built with the library by __graft_entry__.build()
        import subprocess
        r = subprocess.run([drv, "abi"], capture_output=True, text=True, timeout=60)
        assert r.returncode == 0, r.stdout + r.stderr
        assert "fortran 216 library 216" in r.stdout


def _shtns_lib():
    import gfsdc
    lib = ctypes.CDLL(gfsdc.LIBPATH)
    lib.gfsdc_last_error.restype = ctypes.c_char_p
    return lib


def test_shtns_layer_reports_errors_in_record_mode():
    """In mode 1 the SHTns symbols record an error and return (no GPU call is
    made for any of these: the checks come before device work)."""
    lib = _shtns_lib()
    I = ctypes.c_int
    lib.gfsdc_shtns_error_mode(1)
    try:
        lib.gfsdc_shtns_status()                       # clear
        lib.shtns_reset_()
        u = (ctypes.c_double * 8)()
        lib.shtns_sphtor_to_spat_(u, u, u, u)           # before any precompute
        assert lib.gfsdc_shtns_status() != 0
        assert b"precompute" in lib.gfsdc_last_error()
        assert lib.gfsdc_shtns_status() == 0            # status clears once read
        lib.shtns_cos_array_(u)
        assert lib.gfsdc_shtns_status() != 0
        # only the reference's configuration is supported: mmax == lmax, mres 1,
        # SHT_ORTHONORMAL + SHT_NO_CS_PHASE (shtns.f90:110)
        lmax, mmax, mres, norm = I(42), I(40), I(1), I(1024)
        lib.shtns_set_size_(ctypes.byref(lmax), ctypes.byref(mmax), ctypes.byref(mres), ctypes.byref(norm))
        assert lib.gfsdc_shtns_status() != 0
        mmax = I(42)
        lib.shtns_set_size_(ctypes.byref(lmax), ctypes.byref(mmax), ctypes.byref(mres), ctypes.byref(norm))
        assert lib.gfsdc_shtns_status() == 0
        # a regular (non-Gauss) grid is refused before any device work
        typ, lay, eps, nlat, nphi = I(2), I(512), ctypes.c_double(1e-10), I(64), I(128)
        lib.shtns_precompute_(ctypes.byref(typ), ctypes.byref(lay), ctypes.byref(eps), ctypes.byref(nlat),
                              ctypes.byref(nphi))
        assert lib.gfsdc_shtns_status() != 0 and b"Gauss" in lib.gfsdc_last_error()
        nlm = I(0)
        lib.shtns_calc_nlm_(ctypes.byref(nlm), ctypes.byref(lmax), ctypes.byref(mmax), ctypes.byref(mres))
        assert nlm.value == 43 * 44 // 2
    finally:
        lib.gfsdc_shtns_error_mode(0)


def test_shtns_layer_exits_in_default_mode():
    """In the default mode an error ends the process with status 1 and a
    message, as SHTns and the wrapper's stop do (child process, no GPU call)."""
    import subprocess
    import sys
    code = ("import ctypes, sys; sys.path.insert(0, %r); import gfsdc; "
            "lib = ctypes.CDLL(gfsdc.LIBPATH); u = (ctypes.c_double * 8)(); "
            "lib.shtns_gauss_wts_(u); print('not reached')") % os.path.join(ROOT, "gfs-dycore_amd")
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    assert r.returncode == 1, (r.returncode, r.stdout, r.stderr)
    assert "not reached" not in r.stdout and "shtns_gauss_wts" in r.stderr
