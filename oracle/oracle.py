"""ctypes binding of the FP64 CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker. Array conventions follow
the reference (src/shtns.f90, src/spectral_data.f90, src/grid_data.f90):
spectral arrays are complex128 of shape (nfield, ndimspec) (Fortran
(ndimspec, nlev) column-major), grids are float64 (nfield, nlat, nlon)
(Fortran (nlon, nlat, nlev)), latitudes north -> south, state levels
bottom -> top.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_cp = np.ctypeslib.ndpointer(dtype=np.complex128, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


class Cfg(C.Structure):
    _fields_ = [
        ("nlon", C.c_int), ("nlat", C.c_int), ("ntrunc", C.c_int), ("nlev", C.c_int),
        ("ntrac", C.c_int), ("dt", C.c_double),
        ("ak", C.POINTER(C.c_double)), ("bk", C.POINTER(C.c_double)),
        ("explicit_", C.c_int), ("physics", C.c_int), ("massfix", C.c_int), ("ndiss", C.c_int),
        ("efold", C.c_double), ("hdif_fac", C.c_double), ("hdif_fac2", C.c_double),
        ("fshk", C.c_double), ("taustratdamp", C.c_double),
        ("rerth", C.c_double), ("grav", C.c_double), ("omega", C.c_double), ("rd", C.c_double),
        ("rv", C.c_double), ("cp", C.c_double), ("cvap", C.c_double),
        ("simple_test", C.c_int), ("turbflux", C.c_int),
        ("tpat_amp", C.c_double), ("tpat_tau", C.c_double), ("tpat_lscale", C.c_double),
        ("tpat_seed", C.c_int),
    ]


class StochCfg(C.Structure):
    """orc_stoch_cfg (params.f90:140-160 names)"""
    _fields_ = [("vcamp", C.c_double),
                ("svc", C.c_double), ("svc_tau", C.c_double), ("svc_lscale", C.c_double), ("iseed_svc", C.c_int),
                ("sppt", C.c_double), ("sppt_tau", C.c_double), ("sppt_lscale", C.c_double),
                ("iseed_sppt", C.c_int),
                ("shum", C.c_double), ("shum_tau", C.c_double), ("shum_lscale", C.c_double),
                ("iseed_shum", C.c_int),
                ("addnoise", C.c_double), ("addnoise_tau", C.c_double), ("addnoise_lscale", C.c_double),
                ("iseed_addnoise", C.c_int),
                ("addnoise_vfilt", C.c_int), ("addnoise_kenorm", C.c_int), ("addnoise_dissfact", C.c_int),
                ("addnoise_vrtonly", C.c_int)]


def stoch_cfg(**kw):
    c = StochCfg()
    c.addnoise_vrtonly = 1
    for k, v in kw.items():
        setattr(c, k, v)
    return c


def build():
    """Compile liboracle.so in place (gcc)."""
    import subprocess
    subprocess.check_call(["make", "-s", "-C", _HERE, "liboracle.so"])


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        vp = C.c_void_p
        L.orc_sht_create.restype = vp
        L.orc_sht_create.argtypes = [C.c_int, C.c_int, C.c_int]
        L.orc_sht_destroy.argtypes = [vp]
        L.orc_sht_nlm.argtypes = [vp]
        L.orc_sht_griddata.argtypes = [vp, _dp, _dp, _dp, _ip, _ip, _dp, _dp]
        L.orc_sht_legendre.argtypes = [vp, C.c_int, C.c_int, _dp]
        L.orc_spectogrd.argtypes = [vp, _cp, _dp, C.c_int]
        L.orc_grdtospec.argtypes = [vp, _dp, _cp, C.c_int]
        L.orc_getuv.argtypes = [vp, _cp, _cp, _dp, _dp, C.c_int, C.c_double]
        L.orc_getvrtdivspec.argtypes = [vp, _dp, _dp, _cp, _cp, C.c_int, C.c_double]
        L.orc_getgrad.argtypes = [vp, _cp, _dp, _dp, C.c_int, C.c_double]
        L.orc_fft_analysis.argtypes = [C.c_int, C.c_int, _dp, _cp]
        L.orc_fft_synthesis.argtypes = [C.c_int, C.c_int, _cp, _dp]
        L.orc_gauss.argtypes = [C.c_int, _dp, _dp]
        L.orc_cfg_defaults.argtypes = [C.POINTER(Cfg)]
        L.orc_model_create.restype = vp
        L.orc_model_create.argtypes = [C.POINTER(Cfg)]
        L.orc_model_destroy.argtypes = [vp]
        L.orc_set_state.argtypes = [vp, vp, vp, vp, vp, vp]
        L.orc_get_state.argtypes = [vp, vp, vp, vp, vp, vp]
        L.orc_set_pdryini.argtypes = [vp, C.c_double]
        L.orc_get_pdryini.argtypes = [vp]
        L.orc_get_pdryini.restype = C.c_double
        L.orc_set_time.argtypes = [vp, C.c_double]
        L.orc_get_time.argtypes = [vp]
        L.orc_get_time.restype = C.c_double
        L.orc_iau_set.argtypes = [vp, C.c_int, vp, C.c_double, vp]
        L.orc_set_topo.argtypes = [vp, vp]
        L.orc_stoch_init.argtypes = [vp, C.POINTER(StochCfg)]
        L.orc_get_stoch.argtypes = [vp, C.c_int, vp]
        L.orc_set_stoch.argtypes = [vp, C.c_int, vp]
        L.orc_stoch_forcing.argtypes = [vp]
        L.orc_getdyntend.argtypes = [vp, C.c_int, C.c_int, vp, vp, vp, vp, vp]
        L.orc_advance.argtypes = [vp, C.c_int]
        L.orc_hs_phytend.argtypes = [vp, _cp, _cp, _cp]
        L.orc_simple_phytend.argtypes = [vp, _cp, _cp, _cp, _cp]
        L.orc_simple_column.argtypes = [C.c_int, C.c_double, C.c_double, _dp, _dp, _dp, _dp, _dp, _dp,
                                        _dp, _dp, C.c_double, C.POINTER(C.c_double), C.c_int, C.c_int]
        L.orc_simple_sst.argtypes = [C.c_int, C.c_double]
        L.orc_simple_sst.restype = C.c_double
        L.orc_get_grid.argtypes = [vp, C.c_int, _dp]
        L.orc_get_semimp.argtypes = [vp, vp, vp, vp, vp, vp]
        L.orc_get_damp.argtypes = [vp, vp, vp, vp]
        L.orc_step_diag.argtypes = [vp, _dp]
        L.orc_calc_pressdata.argtypes = [C.c_int, C.c_int, _dp, _dp, C.c_double, _dp, _dp, _dp,
                                         _dp, _dp, _dp]
        L.orc_get_pattern.argtypes = [vp, vp, vp, vp, vp]
        L.orc_mlp_nparams.restype = C.c_long
        L.orc_mlp_nparams.argtypes = [C.c_int] * 4
        L.orc_mlp_loss_grad.restype = C.c_double
        L.orc_mlp_loss_grad.argtypes = [C.c_int] * 4 + [_dp, C.c_int, _dp, _dp, vp, vp]
        L.orc_mlp_loss_grad_bf16.restype = C.c_double
        L.orc_mlp_loss_grad_bf16.argtypes = [C.c_int] * 4 + [_dp, C.c_int, _dp, _dp, vp, vp]
        L.orc_mlp_adamw.argtypes = [C.c_int] * 4 + [_dp, _dp, _dp, _dp, C.c_int, C.c_double, C.c_double]
        L.orc_pattern_advance.argtypes = [vp]
        L.orc_vadv_tracers_cols.argtypes = [C.c_int, C.c_long, _dp, _dp, _dp, _dp]
        L.orc_vadv_cols.argtypes = [C.c_int, C.c_long, _dp, _dp, _dp, _dp]
        L.orc_omega_cols.argtypes = [C.c_int, C.c_long] + [_dp] * 12 + [_dp] * 3
        L.orc_presgrad_cols.argtypes = [C.c_int, C.c_long, C.c_double] + [_dp] * 14 + [_dp] * 2
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_get_threads.restype = C.c_int
        L.orc_rng_u64.restype = C.c_uint64
        L.orc_rng_u64.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64]
        L.orc_rng_uniform.restype = C.c_double
        L.orc_rng_uniform.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64]
        _LIB = L
    return _LIB


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def gauss(nlat):
    mu = np.zeros(nlat)
    w = np.zeros(nlat)
    lib().orc_gauss(nlat, mu, w)
    return mu, w


class Sht:
    """Transforms on a Gauss grid with triangular truncation (shtns.f90 API)."""

    def __init__(self, nlon, nlat, ntrunc, _handle=None):
        self.nlon, self.nlat, self.ntrunc = nlon, nlat, ntrunc
        self._own = _handle is None
        self.h = _handle if _handle is not None else lib().orc_sht_create(nlon, nlat, ntrunc)
        self.nlm = lib().orc_sht_nlm(self.h)
        self.lats = np.zeros(nlat)
        self.gauwts = np.zeros(nlat)
        self.areawts = np.zeros(nlat * nlon)
        self.degree = np.zeros(self.nlm, np.int32)
        self.order = np.zeros(self.nlm, np.int32)
        self.lap = np.zeros(self.nlm)
        self.invlap = np.zeros(self.nlm)
        lib().orc_sht_griddata(self.h, self.lats, self.gauwts, self.areawts, self.degree,
                               self.order, self.lap, self.invlap)
        self.areawts = self.areawts.reshape(nlat, nlon)

    def __del__(self):
        if getattr(self, "_own", False) and self.h and _LIB is not None:
            _LIB.orc_sht_destroy(self.h)
            self.h = None

    def legendre(self, m, j):
        out = np.zeros(self.ntrunc + 2 - m)
        lib().orc_sht_legendre(self.h, m, j, out)
        return out

    def _s(self, a):
        a = np.ascontiguousarray(a, dtype=np.complex128)
        return a.reshape(-1, self.nlm), a.shape[:-1]

    def _g(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        return a.reshape(-1, self.nlat, self.nlon), a.shape[:-2]

    def spectogrd(self, spec):
        s, lead = self._s(spec)
        g = np.zeros((s.shape[0], self.nlat, self.nlon))
        lib().orc_spectogrd(self.h, s, g, s.shape[0])
        return g.reshape(lead + (self.nlat, self.nlon))

    def grdtospec(self, grid):
        g, lead = self._g(grid)
        s = np.zeros((g.shape[0], self.nlm), np.complex128)
        lib().orc_grdtospec(self.h, g, s, g.shape[0])
        return s.reshape(lead + (self.nlm,))

    def getuv(self, vrt, div, rsphere):
        v1, lead = self._s(vrt)
        v2, _ = self._s(div)
        u = np.zeros((v1.shape[0], self.nlat, self.nlon))
        v = np.zeros_like(u)
        lib().orc_getuv(self.h, v1, v2, u, v, v1.shape[0], rsphere)
        return u.reshape(lead + u.shape[1:]), v.reshape(lead + v.shape[1:])

    def getvrtdivspec(self, u, v, rsphere):
        g1, lead = self._g(u)
        g2, _ = self._g(v)
        vr = np.zeros((g1.shape[0], self.nlm), np.complex128)
        dv = np.zeros_like(vr)
        lib().orc_getvrtdivspec(self.h, g1, g2, vr, dv, g1.shape[0], rsphere)
        return vr.reshape(lead + (self.nlm,)), dv.reshape(lead + (self.nlm,))

    def getgrad(self, spec, rsphere):
        s, lead = self._s(spec)
        gx = np.zeros((s.shape[0], self.nlat, self.nlon))
        gy = np.zeros_like(gx)
        lib().orc_getgrad(self.h, s, gx, gy, s.shape[0], rsphere)
        return gx.reshape(lead + gx.shape[1:]), gy.reshape(lead + gy.shape[1:])


def cfg_defaults(**kw):
    c = Cfg()
    lib().orc_cfg_defaults(C.byref(c))
    for k, v in kw.items():
        if k in ("ak", "bk"):
            continue
        setattr(c, k, v)
    return c


class Model:
    """Dry/moist dynamical core of src/dyn_run.f90 + src/run.f90 (advance)."""

    GRID = {"ug": 0, "vg": 1, "vrtg": 2, "divg": 3, "virtempg": 4, "dlnpdtg": 5, "etadot": 6,
            "psg": 7, "dlnpsdt": 8, "prs": 9, "pk": 10, "dpk": 11, "alfa": 12, "rlnp": 13,
            "dlnpsdx": 14, "dlnpsdy": 15, "tracerg": 16, "precip": 17, "pwat": 18}

    def __init__(self, nlon, nlat, ntrunc, nlev, ak, bk, dt, ntrac=0, **kw):
        self.ak = np.ascontiguousarray(ak, np.float64)
        self.bk = np.ascontiguousarray(bk, np.float64)
        assert self.ak.size == nlev + 1 and self.bk.size == nlev + 1
        c = cfg_defaults(**kw)
        c.nlon, c.nlat, c.ntrunc, c.nlev, c.ntrac, c.dt = nlon, nlat, ntrunc, nlev, ntrac, dt
        c.ak = self.ak.ctypes.data_as(C.POINTER(C.c_double))
        c.bk = self.bk.ctypes.data_as(C.POINTER(C.c_double))
        self.cfg = c
        self.h = lib().orc_model_create(C.byref(c))
        self.nlon, self.nlat, self.ntrunc, self.nlev, self.ntrac = nlon, nlat, ntrunc, nlev, ntrac
        self.sht = Sht(nlon, nlat, ntrunc)
        self.nlm = self.sht.nlm

    def __del__(self):
        if getattr(self, "h", None) and _LIB is not None:
            _LIB.orc_model_destroy(self.h)
            self.h = None

    def set_state(self, vrt, div, tv, lnps, trac=None):
        a = [np.ascontiguousarray(x, np.complex128) if x is not None else None
             for x in (vrt, div, tv, trac, lnps)]
        lib().orc_set_state(self.h, *[_ptr(x) for x in a])

    def get_state(self):
        s3 = (self.nlev, self.nlm)
        vrt, div, tv = (np.zeros(s3, np.complex128) for _ in range(3))
        trac = np.zeros((max(self.ntrac, 0), self.nlev, self.nlm), np.complex128)
        lnps = np.zeros(self.nlm, np.complex128)
        lib().orc_get_state(self.h, _ptr(vrt), _ptr(div), _ptr(tv),
                            _ptr(trac) if self.ntrac > 0 else None, _ptr(lnps))
        return dict(vrt=vrt, div=div, tv=tv, trac=trac, lnps=lnps)

    @property
    def pdryini(self):
        return lib().orc_get_pdryini(self.h)

    @pdryini.setter
    def pdryini(self, v):
        lib().orc_set_pdryini(self.h, float(v))

    def set_topo(self, topo):
        """spectral orography (m), dyn_init.f90:84"""
        t = np.ascontiguousarray(topo, np.complex128)
        lib().orc_set_topo(self.h, t.ctypes.data)

    def stoch_init(self, **kw):
        """init_stochdata with the params.f90 stochastic settings given as keywords"""
        self._stoch = stoch_cfg(**kw)
        if lib().orc_stoch_init(self.h, C.byref(self._stoch)) != 0:
            raise ValueError("stochastic scheme with amplitude > 0 needs tau > 0 (and addnoise_vrtonly)")

    def stoch(self, which):
        """'svc', 'sppt', 'shum' grids; 'svc_spec', 'sppt_spec', 'shum_spec'; 'addnoise'
        (nlev, nlm) patterns; 'psiforcing' (nlev, nlm), the last step's forcing"""
        w = self.STOCH.index(which)
        if w < 3:
            out = np.zeros((self.nlat, self.nlon))
        elif w < 6:
            out = np.zeros(self.nlm, np.complex128)
        elif w < 8:
            out = np.zeros((self.nlev, self.nlm), np.complex128)
        else:
            out = np.zeros(self.nlev)
        lib().orc_get_stoch(self.h, w, out.ctypes.data)
        return out

    STOCH = ["svc", "sppt", "shum", "svc_spec", "sppt_spec", "shum_spec", "addnoise", "psiforcing",
             "svc_vfact", "sppt_vfact", "shum_vfact", "addnoise_vfact"]

    def set_stoch(self, which, spec):
        """replace the spectral pattern 'svc_spec', 'sppt_spec', 'shum_spec' or 'addnoise'"""
        a = np.ascontiguousarray(spec, np.complex128)
        lib().orc_set_stoch(self.h, self.STOCH.index(which), a.ctypes.data)

    def stoch_forcing(self):
        """getstochforcing on the current state and grid fields"""
        lib().orc_stoch_forcing(self.h)
        return self.stoch("psiforcing")

    @property
    def time(self):
        """model clock (s) at the start of the next step"""
        return lib().orc_get_time(self.h)

    @time.setter
    def time(self, t):
        lib().orc_set_time(self.h, float(t))

    def iau_set(self, fhrs, delthrs, vrt, div, tv, lnps, trac=None):
        """Increments (analysis - first guess) at the hours fhrs: vrt, div, tv
        (nfiles, nlev, nlm), lnps (nfiles, nlm), trac (nfiles, ntrac, nlev, nlm);
        an empty fhrs switches the forcing off."""
        nf = len(fhrs)
        if nf == 0:
            lib().orc_iau_set(self.h, 0, None, 0.0, None)
            return
        parts = [np.asarray(vrt).reshape(nf, -1), np.asarray(div).reshape(nf, -1),
                 np.asarray(tv).reshape(nf, -1)]
        if self.ntrac:
            parts.append(np.asarray(trac).reshape(nf, -1))
        parts.append(np.asarray(lnps).reshape(nf, -1))
        inc = np.ascontiguousarray(np.concatenate(parts, axis=1), np.complex128)
        h = np.ascontiguousarray(fhrs, np.float64)
        if lib().orc_iau_set(self.h, nf, h.ctypes.data, float(delthrs), inc.ctypes.data) != 0:
            raise ValueError("iau times must be two or more, increasing and evenly spaced")

    def dfi(self, ntdfi):
        """Digital filter initialisation of run.f90:58-145: 2 ntdfi steps,
        the state replaced by the dfi_weights(ntdfi) average of the 2 ntdfi + 1
        states of the window and the clock set to the window's middle."""
        w = dfi_weights(ntdfi)
        keys = ("vrt", "div", "tv", "trac", "lnps")
        t0 = self.time
        s = self.get_state()
        acc = {k: w[0] * s[k] for k in keys}
        for nt in range(1, 2 * ntdfi + 1):
            self.advance(1)
            s = self.get_state()
            for k in keys:
                acc[k] = acc[k] + w[nt] * s[k]
        self.set_state(acc["vrt"], acc["div"], acc["tv"], acc["lnps"],
                       acc["trac"] if self.ntrac else None)
        self.time = t0 + ntdfi * self.cfg.dt

    def getdyntend(self, rkstage=1, inverse_only=False):
        s3 = (self.nlev, self.nlm)
        out = dict(dvrt=np.zeros(s3, np.complex128), ddiv=np.zeros(s3, np.complex128),
                   dtv=np.zeros(s3, np.complex128),
                   dtrac=np.zeros((max(self.ntrac, 1), self.nlev, self.nlm), np.complex128),
                   dlnps=np.zeros(self.nlm, np.complex128))
        lib().orc_getdyntend(self.h, rkstage, int(inverse_only), _ptr(out["dvrt"]),
                             _ptr(out["ddiv"]), _ptr(out["dtv"]), _ptr(out["dtrac"]),
                             _ptr(out["dlnps"]))
        return out

    def advance(self, nsteps=1):
        lib().orc_advance(self.h, nsteps)

    def hs_phytend(self):
        s3 = (self.nlev, self.nlm)
        a, b, c = (np.zeros(s3, np.complex128) for _ in range(3))
        lib().orc_hs_phytend(self.h, a, b, c)
        return dict(dvrt=a, ddiv=b, dtv=c)

    def simple_phytend(self):
        s3 = (self.nlev, self.nlm)
        a, b, c = (np.zeros(s3, np.complex128) for _ in range(3))
        q = np.zeros((max(self.ntrac, 1), self.nlev, self.nlm), np.complex128)
        lib().orc_simple_phytend(self.h, a, b, c, q)
        return dict(dvrt=a, ddiv=b, dtv=c, dtrac=q)

    def pattern(self):
        """AR(1) temperature pattern: dict(spec, varspectrum, phi, vfact)"""
        spec = np.zeros(self.nlm, np.complex128)
        var = np.zeros(self.nlm)
        phi = C.c_double(0.0)
        vf = np.zeros(self.nlev)
        lib().orc_get_pattern(self.h, _ptr(spec), _ptr(var), C.byref(phi), _ptr(vf))
        return dict(spec=spec, varspectrum=var, phi=phi.value, vfact=vf)

    def pattern_advance(self):
        lib().orc_pattern_advance(self.h)

    def grid(self, name):
        n2 = self.nlat * self.nlon
        nk = {"etadot": self.nlev + 1, "pk": self.nlev + 1, "psg": 1, "dlnpsdt": 1,
              "dlnpsdx": 1, "dlnpsdy": 1, "tracerg": self.nlev * self.ntrac, "precip": 1,
              "pwat": 1}.get(name, self.nlev)
        out = np.zeros(nk * n2)
        lib().orc_get_grid(self.h, self.GRID[name], out)
        return out.reshape(nk, self.nlat, self.nlon)

    def semimp(self):
        n = self.nlev
        A = np.zeros(n * n); B = np.zeros(n * n); sv = np.zeros(n); tor = np.zeros(n)
        D = np.zeros(n * n * (self.ntrunc + 1) * 3)
        lib().orc_get_semimp(self.h, _ptr(A), _ptr(B), _ptr(sv), _ptr(tor), _ptr(D))
        # column-major (Fortran) -> numpy [row, col]
        return dict(amhyb=A.reshape(n, n).T.copy(), bmhyb=B.reshape(n, n).T.copy(), svhyb=sv,
                    tor_hyb=tor,
                    d_hyb_m=D.reshape(3, self.ntrunc + 1, n, n).transpose(0, 1, 3, 2).copy())

    def damp(self):
        d = np.zeros(self.nlm); f = np.zeros(self.nlev); p = np.zeros(self.nlev)
        lib().orc_get_damp(self.h, _ptr(d), _ptr(f), _ptr(p))
        return dict(disspec=d, diff_prof=f, dmp_prof=p)

    def step_diag(self):
        out = np.zeros(4)
        lib().orc_step_diag(self.h, out)
        return dict(spdmax=out[0], psmin=out[1], psmax=out[2], pstend=out[3])


def dfi_weights(ntdfi):
    """set_dfi_wts (run.f90:567-587): Lanczos-windowed sinc over the 2 ntdfi + 1
    states of the window, sin(wx)/wx sin(sx)/sx with sx = pi (k - N) / N,
    wx = pi (k - N) / (N + 1), 1 at k = N, normalised to sum 1. pi is the
    default-real acos(-1.)."""
    pi = float(np.float32(np.pi))
    k = np.arange(2 * ntdfi + 1, dtype=np.float64) - ntdfi
    w = np.ones(2 * ntdfi + 1)
    nz = k != 0
    sx = pi * k[nz] / ntdfi
    wx = pi * k[nz] / (ntdfi + 1)
    w[nz] = np.sin(wx) / wx * np.sin(sx) / sx
    return w / w.sum()


def calc_pressdata(ak, bk, rk, lnpsg):
    ak = np.ascontiguousarray(ak, np.float64)
    bk = np.ascontiguousarray(bk, np.float64)
    lnpsg = np.ascontiguousarray(lnpsg, np.float64).ravel()
    nlev, ncol = ak.size - 1, lnpsg.size
    pk = np.zeros((nlev + 1, ncol)); dpk = np.zeros((nlev, ncol)); alfa = np.zeros((nlev, ncol))
    rlnp = np.zeros((nlev, ncol)); prs = np.zeros((nlev, ncol))
    lib().orc_calc_pressdata(nlev, ncol, ak, bk, rk, lnpsg, pk, dpk, alfa, rlnp, prs)
    return dict(pk=pk, dpk=dpk, alfa=alfa, rlnp=rlnp, prs=prs)


def vadv(datag, etadot, dpk):
    """getvadv on columns: datag (nlev, ncol) bottom -> top, etadot (nlev+1, ncol)
    and dpk (nlev, ncol) top -> bottom."""
    datag, etadot, dpk = (np.ascontiguousarray(x, np.float64) for x in (datag, etadot, dpk))
    nl, n = datag.shape
    out = np.zeros_like(datag)
    lib().orc_vadv_cols(nl, n, datag, etadot, dpk, out)
    return out


def vadv_tracers(datag, etadot, dpk):
    """getvadv_tracers on columns: datag (nlev, ncol) bottom -> top, etadot
    (nlev+1, ncol) and dpk (nlev, ncol) top -> bottom."""
    datag, etadot, dpk = (np.ascontiguousarray(x, np.float64) for x in (datag, etadot, dpk))
    nl, n = datag.shape
    out = np.zeros_like(datag)
    lib().orc_vadv_tracers_cols(nl, n, datag, etadot, dpk, out)
    return out


def _press_cols(ak, bk, lnpsg, rk):
    ak = np.ascontiguousarray(ak, np.float64)
    bk = np.ascontiguousarray(bk, np.float64)
    p = calc_pressdata(ak, bk, rk, lnpsg)
    dbk = np.diff(bk)
    ck = ak[1:] * bk[:-1] - ak[:-1] * bk[1:]
    return ak, bk, dbk, ck, np.exp(np.ascontiguousarray(lnpsg, np.float64).ravel()), p


def omega_cols(ak, bk, lnpsg, ug, vg, divg, dlnpsdx, dlnpsdy, rk=287.05 / 1004.6):
    """getomega on columns: ug, vg, divg (nlev, ncol) bottom -> top; lnpsg,
    dlnpsdx, dlnpsdy (ncol). Pressure fields come from calc_pressdata on the
    same columns. Returns dlnpsdt (ncol), etadot (nlev+1, ncol) top -> bottom
    and dlnpdtg = omega/p (nlev, ncol) bottom -> top."""
    ak, bk, dbk, ck, ps, p = _press_cols(ak, bk, lnpsg, rk)
    ug, vg, divg, dx, dy = (np.ascontiguousarray(x, np.float64) for x in (ug, vg, divg, dlnpsdx, dlnpsdy))
    nl, n = ug.shape
    dlnpsdt = np.zeros(n); etadot = np.zeros((nl + 1, n)); omg = np.zeros((nl, n))
    lib().orc_omega_cols(nl, n, bk, dbk, ck, ps, p["dpk"], p["alfa"], p["rlnp"], ug, vg, divg, dx, dy,
                         dlnpsdt, etadot, omg)
    return dlnpsdt, etadot, omg


def presgrad_cols(ak, bk, lnpsg, tvg, dtvdx, dtvdy, dphisdx, dphisdy, dlnpsdx, dlnpsdy, rd=287.05,
                  rk=287.05 / 1004.6):
    """getpresgrad on columns: tvg, dtvdx, dtvdy (nlev, ncol) bottom -> top;
    lnpsg, dphisdx/y, dlnpsdx/y (ncol). Returns prsgx, prsgy (nlev, ncol)
    bottom -> top."""
    ak, bk, dbk, ck, ps, p = _press_cols(ak, bk, lnpsg, rk)
    arrs = [np.ascontiguousarray(x, np.float64) for x in (tvg, dtvdx, dtvdy, dphisdx, dphisdy, dlnpsdx, dlnpsdy)]
    nl, n = arrs[0].shape
    px = np.zeros((nl, n)); py = np.zeros((nl, n))
    lib().orc_presgrad_cols(nl, n, rd, bk, dbk, ps, p["pk"], p["dpk"], p["alfa"], p["rlnp"], *arrs, px, py)
    return px, py


class Mlp:
    """FP64 column MLP: nin -> hidden x nlayers (GELU) -> nout; flat
    parameters per layer W[K][N] then b[N]."""

    def __init__(self, nin, nout, hidden, nlayers):
        self.dims = (nin, nout, hidden, nlayers)
        self.nparams = lib().orc_mlp_nparams(*self.dims)

    def loss_grad(self, params, x, t, grad=True, bf16=False):
        """loss, parameter gradient and outputs; bf16=True rounds inputs,
        weights, hidden activations and back-propagated gradients to bfloat16
        as a bf16 matrix-core network stores them (sums in FP64)"""
        params = np.ascontiguousarray(params, np.float64)
        x = np.ascontiguousarray(x, np.float64)
        t = np.ascontiguousarray(t, np.float64)
        g = np.zeros(self.nparams) if grad else None
        y = np.zeros((x.shape[0], self.dims[1]))
        f = lib().orc_mlp_loss_grad_bf16 if bf16 else lib().orc_mlp_loss_grad
        loss = f(*self.dims, params, x.shape[0], x, t, _ptr(g), _ptr(y))
        return loss, g, y

    def adamw(self, params, m, v, grad, step, lr, wd):
        for a in (params, m, v):
            assert a.dtype == np.float64 and a.flags.c_contiguous
        lib().orc_mlp_adamw(*self.dims, params, m, v, np.ascontiguousarray(grad, np.float64), step, lr, wd)


def simple_column(dtime, lat, t, q, u, v, pmid, pint, pdel, rpdel, ps, sst_mode=0, turbflux=True):
    """simple_physics_v5_1.f90 on one column (top -> bottom); returns the
    updated (t, q, u, v) and the precipitation rate (m/s)."""
    t, q, u, v = (np.array(x, np.float64) for x in (t, q, u, v))
    pmid, pint, pdel, rpdel = (np.ascontiguousarray(x, np.float64) for x in (pmid, pint, pdel, rpdel))
    pr = C.c_double(0.0)
    lib().orc_simple_column(t.size, dtime, lat, t, q, u, v, pmid, pint, pdel, rpdel, ps, C.byref(pr),
                            int(sst_mode), int(bool(turbflux)))
    return t, q, u, v, pr.value


def simple_sst(sst_mode, lat):
    return lib().orc_simple_sst(int(sst_mode), float(lat))


def set_threads(n):
    """OpenMP threads of the oracle's loops; returns the previous count"""
    prev = lib().orc_get_threads()
    lib().orc_set_threads(int(n))
    return prev


def rng_uniform(seed, stream, index):
    return lib().orc_rng_uniform(seed, stream, index)
