"""The drop-in boundary on the GPU.

1. The eleven SHTns symbols of libgfsdc driven exactly as src/shtns.f90 drives
   SHTns: shtns_init's call sequence (shtns.f90:109-126) and module data, then
   each wrapper routine with its own pre- and post-scaling --
   getuv's S = invlap a vrt, T = invlap a div (shtns.f90:259-261), getvrtdivspec's
   lap / a (:310), getgrad's S = 0 and 1 / a (:339-351). The wrapper is restated
   here in Python over the symbols and compared with the FP64 oracle (1e-11).
   The symbols are also called from several host threads at once, as the
   reference calls the wrapper inside OpenMP loops over levels.
2. The Fortran host layer: gfsdc_hs_driver (module shtns, gfsdc_dycore and a
   run.f90-shaped time loop) runs three T42L20 Held-Suarez steps; its shtns
   round trip and solid-body getuv must be exact to FP32 rounding and its
   per-step diagnostics (run.f90:159-163) must equal the oracle's.
"""
import ctypes as C
import os
import re
import subprocess
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import cases
import levels
import oracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
A = float(np.float32(6.3712e6))
SHT_ORTHONORMAL, SHT_NO_CS_PHASE, SHT_GAUSS_FLY, SHT_PHI_CONTIGUOUS = 0, 1024, 6, 512


def _ip(v):
    return C.byref(C.c_int(v))


class RefWrapper:
    """src/shtns.f90's wrapper restated over the SHTns symbols (FP64 arrays)."""

    def __init__(self, nlon, nlat, ntrunc):
        import gfsdc
        L = C.CDLL(gfsdc.LIBPATH)
        L.gfsdc_last_error.restype = C.c_char_p
        L.gfsdc_shtns_error_mode(1)
        self.L = L
        self.nlon, self.nlat, T = nlon, nlat, ntrunc
        # shtns.f90:108-112
        L.shtns_use_threads_(_ip(1))
        L.shtns_set_size_(_ip(T), _ip(T), _ip(1), _ip(SHT_ORTHONORMAL + SHT_NO_CS_PHASE))
        L.shtns_precompute_(_ip(SHT_GAUSS_FLY), _ip(SHT_PHI_CONTIGUOUS), C.byref(C.c_double(1e-10)),
                            _ip(nlat), _ip(nlon))
        nlm = C.c_int(0)
        L.shtns_calc_nlm_(C.byref(nlm), _ip(T), _ip(T), _ip(1))
        self.ok()
        self.nlm = nlm.value
        assert self.nlm == (T + 1) * (T + 2) // 2                       # shtns.f90:116
        # shtns.f90:125-139
        lats1 = np.zeros(nlat)
        w1 = np.zeros(nlat // 2)
        L.shtns_cos_array_(lats1.ctypes.data_as(C.c_void_p))
        L.shtns_gauss_wts_(w1.ctypes.data_as(C.c_void_p))
        self.ok()
        self.lats = np.arcsin(lats1)
        self.gauwts = np.concatenate([w1, w1[::-1]])
        aw = np.broadcast_to(self.gauwts[:, None], (nlat, nlon))
        self.areawts = aw / aw.sum()
        # shtns.f90:149-155
        self.degree = np.concatenate([np.arange(m, T + 1) for m in range(T + 1)])
        self.order = np.concatenate([np.full(T + 1 - m, m) for m in range(T + 1)])
        self.lap = -self.degree * (self.degree + 1.0)
        self.invlap = np.zeros(self.nlm)
        self.invlap[1:] = 1.0 / self.lap[1:]

    def ok(self):
        st = self.L.gfsdc_shtns_status()
        assert st == 0, self.L.gfsdc_last_error()

    def _g(self):
        return np.zeros((self.nlat, self.nlon))

    def _s(self):
        return np.zeros(self.nlm, np.complex128)

    @staticmethod
    def _v(a):
        return a.ctypes.data_as(C.c_void_p)

    def grdtospec(self, g):                                            # shtns.f90:171-201
        g = np.ascontiguousarray(g, np.float64)
        s = self._s()
        self.L.shtns_spat_to_sh_(self._v(g), self._v(s))
        self.ok()
        return s

    def spectogrd(self, s):                                            # shtns.f90:203-231
        s = np.ascontiguousarray(s, np.complex128)
        g = self._g()
        self.L.shtns_sh_to_spat_(self._v(s), self._v(g))
        self.ok()
        return g

    def getuv(self, vrt, div, rs):                                     # shtns.f90:233-273
        S = self.invlap * rs * vrt
        T = self.invlap * rs * div
        u, v = self._g(), self._g()
        self.L.shtns_sphtor_to_spat_(self._v(S), self._v(T), self._v(u), self._v(v))
        self.ok()
        return u, v

    def getvrtdivspec(self, u, v, rs):                                 # shtns.f90:275-311
        u = np.ascontiguousarray(u, np.float64)
        v = np.ascontiguousarray(v, np.float64)
        S, T = self._s(), self._s()
        self.L.shtns_spat_to_sphtor_(self._v(u), self._v(v), self._v(S), self._v(T))
        self.ok()
        return (self.lap / rs) * S, (self.lap / rs) * T

    def getgrad(self, f, rs):                                          # shtns.f90:313-352
        z = self._s()
        f = np.ascontiguousarray(f, np.complex128)
        u, v = self._g(), self._g()
        self.L.shtns_sphtor_to_spat_(self._v(z), self._v(f), self._v(u), self._v(v))
        self.ok()
        return u / rs, v / rs

    def close(self):
        self.L.shtns_reset_()
        self.ok()
        self.L.gfsdc_shtns_error_mode(0)


def rand_spec(degree, order, seed, nf=1):
    rng = np.random.default_rng(seed)
    n = np.maximum(degree, 1).astype(float)
    q = n ** -1.5 * np.exp(2j * np.pi * rng.random((nf, degree.size)))
    q[:, order == 0] = q[:, order == 0].real
    return q


def rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


@pytest.mark.parametrize("nlon,nlat,T", [(128, 64, 42), (384, 190, 126)])
def test_shtns_symbols_as_the_reference_wrapper_calls_them(nlon, nlat, T):
    w = RefWrapper(nlon, nlat, T)
    o = O.Sht(nlon, nlat, T)
    try:
        assert np.abs(w.lats - o.lats).max() < 1e-14                   # N -> S
        assert np.abs(w.areawts - o.areawts.reshape(nlat, nlon)).max() < 1e-17
        assert np.array_equal(w.degree, o.degree) and np.array_equal(w.order, o.order)
        qd = rand_spec(o.degree, o.order, 1, 3)
        q = qd[:2].copy()
        q[0] *= 3e-5                                                   # vorticity-sized
        d = qd[2:3] * 1e-6                                             # divergence-sized
        g = o.spectogrd(q[:1])[0]
        assert rel(w.spectogrd(q[0]), g) < 1e-11
        assert rel(w.grdtospec(g), o.grdtospec(g[None])[0]) < 1e-11
        u, v = w.getuv(q[0], d[0], A)
        uo, vo = o.getuv(q[:1], d[:1], A)
        assert rel(u, uo[0]) < 1e-11 and rel(v, vo[0]) < 1e-11
        vr, dv = w.getvrtdivspec(uo[0], vo[0], A)
        vro, dvo = o.getvrtdivspec(uo, vo, A)
        assert rel(vr, vro[0]) < 1e-11 and rel(dv, dvo[0]) < 1e-11
        # and the round trip recovers vorticity and divergence (n >= 1)
        assert rel(vr[1:], q[0][1:]) < 1e-11 and rel(dv[1:], d[0][1:]) < 1e-10
        gx, gy = w.getgrad(q[1], A)
        gxo, gyo = o.getgrad(q[1:], A)
        assert rel(gx, gxo[0]) < 1e-11 and rel(gy, gyo[0]) < 1e-11
    finally:
        w.close()


def test_shtns_symbols_from_concurrent_host_threads():
    """The reference calls the wrapper inside OpenMP loops over levels
    (dyn_run.f90:101,220): concurrent calls on the shared configuration give
    the same fields as serial ones."""
    nlon, nlat, T = 128, 64, 42
    w = RefWrapper(nlon, nlat, T)
    try:
        qs = rand_spec(w.degree, w.order, 7, 16)
        serial = [w.spectogrd(q) for q in qs]
        with ThreadPoolExecutor(8) as ex:
            par = list(ex.map(w.spectogrd, qs))
            back = list(ex.map(w.grdtospec, par))
        for a, b in zip(par, serial):
            assert np.array_equal(a, b)
        for a, q in zip(back, qs):
            assert rel(a, q) < 1e-12
    finally:
        w.close()


def test_fortran_host_layer_hs_driver():
    drv = os.path.join(ROOT, "gfs-dycore_amd", "gfsdc_hs_driver")
    assert os.path.exists(drv), "build the Fortran layer first (make -C gfs-dycore_amd)"
    print("running gfsdc_hs_driver 3 T42L20 fp64", flush=True)
    r = subprocess.run([drv, "3", "T42L20", "fp64"], capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stdout + r.stderr
    out = r.stdout
    rt = float(re.search(r"round trip max error =\s*(\S+)", out).group(1))
    sb = float(re.search(r"solid-body max error =\s*(\S+)", out).group(1))
    assert rt < 1e-6 and sb < 1e-6, (rt, sb)          # FP32 arrays (kinds.f90 r_kind)
    rows = re.findall(r"spdmax =\s*(\S+), min/max ps =\s*(\S+)/\s*(\S+), pstend = (\S+),", out)
    assert len(rows) == 3, out
    # the same three steps in the oracle from the same seeded state
    ak, bk = levels.sigma_levels(20)
    o = O.Model(128, 64, 42, 20, ak, bk, 1200.0)
    vrt, div, tv, lnps = cases.rest_state(o.sht.degree, o.sht.order, 20, seed=0)
    o.set_state(vrt[0], div[0], tv[0], lnps[0])
    for row in rows:
        o.advance(1)
        d = o.step_diag()
        spd, pmin, pmax, pst = (float(x) for x in row)
        assert abs(spd - d["spdmax"]) <= 6e-4 + 1e-6 * d["spdmax"], (spd, d)
        assert abs(pmin - d["psmin"] / 100) <= 6e-3 and abs(pmax - d["psmax"] / 100) <= 6e-3, (row, d)
        assert abs(pst - d["pstend"]) <= 6e-4 + 1e-6 * abs(d["pstend"]), (pst, d)


def test_fortran_host_layer_dfi():
    """run.f90's digital filter window through the Fortran host (dyc_dfi), then
    three steps from the window's middle: diagnostics equal to the oracle's"""
    drv = os.path.join(ROOT, "gfs-dycore_amd", "gfsdc_hs_driver")
    print("running gfsdc_hs_driver 3 T42L20 fp64 2", flush=True)
    r = subprocess.run([drv, "3", "T42L20", "fp64", "2"], capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stdout + r.stderr
    m = re.search(r"done with dfi loop, restarting at t = (\S+) hrs", r.stdout)
    assert m and abs(float(m.group(1)) - 2 * 1200.0 / 3600.0) < 1e-3, r.stdout
    rows = re.findall(r"t = (\S+) hrs, spdmax =\s*(\S+), min/max ps =\s*(\S+)/\s*(\S+), pstend = (\S+),", r.stdout)
    assert len(rows) == 3, r.stdout
    ak, bk = levels.sigma_levels(20)
    o = O.Model(128, 64, 42, 20, ak, bk, 1200.0)
    vrt, div, tv, lnps = cases.rest_state(o.sht.degree, o.sht.order, 20, seed=0)
    o.set_state(vrt[0], div[0], tv[0], lnps[0])
    o.dfi(2)
    for row in rows:
        o.advance(1)
        d = o.step_diag()
        th, spd, pmin, pmax, pst = (float(x) for x in row)
        assert abs(th - o.time / 3600.0) < 1e-3
        assert abs(spd - d["spdmax"]) <= 6e-4 + 1e-6 * d["spdmax"], (spd, d)
        assert abs(pmin - d["psmin"] / 100) <= 6e-3 and abs(pmax - d["psmax"] / 100) <= 6e-3, (row, d)
        assert abs(pst - d["pstend"]) <= 6e-4 + 1e-6 * abs(d["pstend"]), (pst, d)
