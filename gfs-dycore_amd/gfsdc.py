"""Python host binding of libgfsdc (ctypes over the C ABI in include/gfsdc.h).

Mirrors the public interface of the reference's transform module
(src/shtns.f90: spectogrd, grdtospec, getuv, getvrtdivspec, getgrad and the
module data lats/gauwts/areawts/degree/order/lap/invlap) and of the dynamics
call sequence (src/dyn_run.f90 getdyntend, src/run.f90 advance) so tests read
like the reference's own usage. Arrays are numpy, in the reference layout:
spectral complex (nfield, ndimspec), grid real (nfield, nlat, nlon).

The HIP library is required: importing this module raises if libgfsdc.so is
missing, and every call raises GfsdcError on a non-zero status.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(_HERE, "libgfsdc.so")

FP64, FP32, BF16X3 = 0, 1, 2
HOST, DEVICE = 0, 1
DT_F64, DT_F32 = 0, 1
PHYS_NONE, PHYS_HS, PHYS_SIMPLE = 0, 1, 2


class GfsdcError(RuntimeError):
    pass


class Cfg(C.Structure):
    _fields_ = [
        ("nlon", C.c_int), ("nlat", C.c_int), ("ntrunc", C.c_int), ("nlev", C.c_int),
        ("ntrac", C.c_int), ("nmem", C.c_int), ("precision", C.c_int), ("dt", C.c_double),
        ("ak", C.POINTER(C.c_double)), ("bk", C.POINTER(C.c_double)),
        ("explicit_", C.c_int), ("physics", C.c_int), ("massfix", C.c_int), ("ndiss", C.c_int),
        ("efold", C.c_double), ("hdif_fac", C.c_double), ("hdif_fac2", C.c_double),
        ("fshk", C.c_double), ("taustratdamp", C.c_double),
        ("rerth", C.c_double), ("grav", C.c_double), ("omega", C.c_double), ("rd", C.c_double),
        ("rv", C.c_double), ("cp", C.c_double), ("cvap", C.c_double), ("device", C.c_int),
        ("simple_test", C.c_int), ("turbflux", C.c_int),
        ("tpat_amp", C.c_double), ("tpat_tau", C.c_double), ("tpat_lscale", C.c_double),
        ("tpat_seed", C.c_int),
    ]


class StochCfg(C.Structure):
    """gfsdc_stoch_cfg (params.f90:140-160 names)"""
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


SYMBOLS = [
    "shtns_use_threads_", "shtns_set_size_", "shtns_precompute_", "shtns_calc_nlm_",
    "shtns_cos_array_", "shtns_gauss_wts_", "shtns_spat_to_sh_", "shtns_sh_to_spat_",
    "shtns_sphtor_to_spat_", "shtns_spat_to_sphtor_", "shtns_reset_",
    "gfsdc_cfg_defaults", "gfsdc_create", "gfsdc_destroy", "gfsdc_last_error", "gfsdc_nlm",
    "gfsdc_griddata", "gfsdc_spectogrd", "gfsdc_spectogrd_noalloc", "gfsdc_grdtospec", "gfsdc_getuv",
    "gfsdc_getvrtdivspec", "gfsdc_getgrad", "gfsdc_set_state", "gfsdc_get_state",
    "gfsdc_set_topo", "gfsdc_set_pdryini", "gfsdc_get_pdryini", "gfsdc_getdyntend",
    "gfsdc_advance", "gfsdc_hs_phytend", "gfsdc_hs_drag_levels", "gfsdc_simple_phytend", "gfsdc_get_pattern", "gfsdc_autotune", "gfsdc_bench_transforms", "gfsdc_bench_copy", "gfsdc_mlp_init", "gfsdc_mlp_collect",
    "gfsdc_mlp_normalize", "gfsdc_mlp_train", "gfsdc_mlp_eval", "gfsdc_mlp_enable", "gfsdc_mlp_judge_stats",
    "gfsdc_mlp_nparams", "gfsdc_mlp_params", "gfsdc_mlp_norm", "gfsdc_mlp_forward", "gfsdc_mlp_grad", "gfsdc_mlp_adamw",
    "gfsdc_get_grid", "gfsdc_get_semimp",
    "gfsdc_get_damp", "gfsdc_step_diag", "gfsdc_init_rest_state", "gfsdc_set_time",
    "gfsdc_get_time", "gfsdc_iau_set", "gfsdc_dfi", "gfsdc_stoch_defaults", "gfsdc_stoch_init", "gfsdc_get_stoch", "gfsdc_checkpoint_size", "gfsdc_checkpoint_save",
    "gfsdc_checkpoint_load", "gfsdc_sync",
    "gfsdc_stream", "gfsdc_profile_step",
]

_LIB = None


def load():
    """Load libgfsdc.so (raises if it has not been built)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIBPATH):
        raise GfsdcError(f"{LIBPATH} is missing: build it with `make -C {_HERE}` "
                         "(or __graft_entry__.build())")
    L = C.CDLL(LIBPATH)
    vp, i, d = C.c_void_p, C.c_int, C.c_double
    L.gfsdc_last_error.restype = C.c_char_p
    L.gfsdc_cfg_defaults.argtypes = [C.POINTER(Cfg)]
    L.gfsdc_create.argtypes = [C.POINTER(Cfg), C.POINTER(vp)]
    L.gfsdc_destroy.argtypes = [vp]
    L.gfsdc_nlm.argtypes = [vp]
    L.gfsdc_griddata.argtypes = [vp] + [vp] * 7
    for n in ("gfsdc_spectogrd", "gfsdc_grdtospec"):
        getattr(L, n).argtypes = [vp, vp, vp, i, i, i]
    L.gfsdc_getuv.argtypes = [vp, vp, vp, vp, vp, i, d, i, i]
    L.gfsdc_getvrtdivspec.argtypes = [vp, vp, vp, vp, vp, i, d, i, i]
    L.gfsdc_getgrad.argtypes = [vp, vp, vp, vp, i, d, i, i]
    L.gfsdc_set_state.argtypes = [vp, vp, vp, vp, vp, vp, i, i]
    L.gfsdc_get_state.argtypes = [vp, vp, vp, vp, vp, vp, i, i]
    L.gfsdc_set_topo.argtypes = [vp, vp, i, i]
    L.gfsdc_set_pdryini.argtypes = [vp, vp]
    L.gfsdc_get_pdryini.argtypes = [vp, vp]
    L.gfsdc_getdyntend.argtypes = [vp, i, i, vp, vp, vp, vp, vp, i, i]
    L.gfsdc_advance.argtypes = [vp, i, i]
    L.gfsdc_hs_phytend.argtypes = [vp, vp, vp, vp, i, i]
    L.gfsdc_hs_drag_levels.argtypes = [vp, C.POINTER(i)]
    L.gfsdc_simple_phytend.argtypes = [vp, vp, vp, vp, vp, i, i]
    L.gfsdc_get_pattern.argtypes = [vp, vp, i, i]
    L.gfsdc_autotune.argtypes = [vp, vp, C.POINTER(i)]
    L.gfsdc_bench_transforms.argtypes = [vp, i, i, i, C.POINTER(d)]
    L.gfsdc_bench_copy.argtypes = [vp, C.c_long, i, C.POINTER(d)]
    lg = C.c_long
    L.gfsdc_mlp_init.argtypes = [vp, i, i, i, lg, C.c_ulonglong]
    L.gfsdc_mlp_collect.argtypes = [vp, i, i, C.POINTER(lg)]
    L.gfsdc_mlp_normalize.argtypes = [vp]
    L.gfsdc_mlp_train.argtypes = [vp, i, i, d, d, lg, vp]
    L.gfsdc_mlp_eval.argtypes = [vp, lg, lg, C.POINTER(d), C.POINTER(d)]
    L.gfsdc_mlp_enable.argtypes = [vp, i, i, d]
    L.gfsdc_mlp_judge_stats.argtypes = [vp, vp]
    L.gfsdc_mlp_nparams.argtypes = [vp]
    L.gfsdc_mlp_nparams.restype = lg
    L.gfsdc_mlp_params.argtypes = [vp, vp, vp]
    L.gfsdc_mlp_norm.argtypes = [vp, vp, vp]
    L.gfsdc_mlp_forward.argtypes = [vp, vp, lg, vp]
    L.gfsdc_mlp_grad.argtypes = [vp, vp, vp, i, vp, C.POINTER(d)]
    L.gfsdc_mlp_adamw.argtypes = [vp, vp, d, d]
    L.gfsdc_get_grid.argtypes = [vp, i, vp, i, i]
    L.gfsdc_get_semimp.argtypes = [vp, vp, vp, vp, vp, vp]
    L.gfsdc_get_damp.argtypes = [vp, vp, vp, vp]
    L.gfsdc_step_diag.argtypes = [vp, vp]
    L.gfsdc_init_rest_state.argtypes = [vp, d, d, d, i, i, C.c_ulonglong]
    L.gfsdc_set_time.argtypes = [vp, d]
    L.gfsdc_get_time.argtypes = [vp, C.POINTER(d)]
    L.gfsdc_iau_set.argtypes = [vp, i, vp, d, vp, vp, vp, vp, vp, i, i]
    L.gfsdc_dfi.argtypes = [vp, i, i]
    L.gfsdc_stoch_defaults.argtypes = [C.POINTER(StochCfg)]
    L.gfsdc_stoch_init.argtypes = [vp, C.POINTER(StochCfg)]
    L.gfsdc_get_stoch.argtypes = [vp, i, vp, i, i]
    L.gfsdc_checkpoint_size.argtypes = [vp, C.POINTER(C.c_size_t)]
    L.gfsdc_checkpoint_save.argtypes = [vp, vp, C.c_size_t, i]
    L.gfsdc_checkpoint_load.argtypes = [vp, vp, C.c_size_t, i]
    L.gfsdc_sync.argtypes = [vp]
    L.gfsdc_stream.argtypes = [vp]
    L.gfsdc_stream.restype = vp
    L.gfsdc_profile_step.argtypes = [vp, vp, vp, i]
    _LIB = L
    return L


def _check(rc):
    if rc != 0:
        raise GfsdcError(f"gfsdc error {rc}: {load().gfsdc_last_error().decode()}")


def _p(a):
    return None if a is None else C.c_void_p(a.ctypes.data)


def cfg_defaults():
    c = Cfg()
    load().gfsdc_cfg_defaults(C.byref(c))
    return c


class Context:
    """One gfsdc context: a transform plan (and, with nlev > 0, a dycore state)."""

    GRID = {"ug": 0, "vg": 1, "vrtg": 2, "divg": 3, "virtempg": 4, "dlnpdtg": 5, "etadot": 6,
            "psg": 7, "dlnpsdt": 8, "prs": 9, "dlnpsdx": 14, "dlnpsdy": 15, "tracerg": 16,
            "precip": 17, "pwat": 18}

    def __init__(self, nlon, nlat, ntrunc, precision=FP64, nlev=0, ak=None, bk=None, dt=0.0,
                 ntrac=0, nmem=1, **kw):
        L = load()
        c = cfg_defaults()
        c.nlon, c.nlat, c.ntrunc, c.nlev, c.ntrac, c.nmem = nlon, nlat, ntrunc, nlev, ntrac, nmem
        c.precision, c.dt = precision, dt
        for k, v in kw.items():
            setattr(c, k, v)
        self._ak = self._bk = None
        if nlev > 0:
            self._ak = np.ascontiguousarray(ak, np.float64)
            self._bk = np.ascontiguousarray(bk, np.float64)
            assert self._ak.size == nlev + 1 and self._bk.size == nlev + 1
            c.ak = self._ak.ctypes.data_as(C.POINTER(C.c_double))
            c.bk = self._bk.ctypes.data_as(C.POINTER(C.c_double))
        self.cfg = c
        h = C.c_void_p()
        _check(L.gfsdc_create(C.byref(c), C.byref(h)))
        self.h = h
        self.nlon, self.nlat, self.ntrunc = nlon, nlat, ntrunc
        self.nlev, self.ntrac, self.nmem, self.precision = nlev, ntrac, nmem, precision
        self.nlm = L.gfsdc_nlm(self.h)
        self.lats = np.zeros(nlat)
        self.gauwts = np.zeros(nlat)
        self.areawts = np.zeros((nlat, nlon))
        self.degree = np.zeros(self.nlm, np.int32)
        self.order = np.zeros(self.nlm, np.int32)
        self.lap = np.zeros(self.nlm)
        self.invlap = np.zeros(self.nlm)
        _check(L.gfsdc_griddata(self.h, _p(self.lats), _p(self.gauwts), _p(self.areawts),
                                _p(self.degree), _p(self.order), _p(self.lap), _p(self.invlap)))

    def close(self):
        if getattr(self, "h", None) is not None and _LIB is not None:
            _LIB.gfsdc_destroy(self.h)
            self.h = None

    __del__ = close

    # -- dtype plumbing: host numpy arrays in FP64 or FP32 ------------------
    @staticmethod
    def _dt(a):
        return DT_F32 if a.dtype in (np.float32, np.complex64) else DT_F64

    def _spec_in(self, a, dtype):
        a = np.ascontiguousarray(a, dtype=np.complex64 if dtype == DT_F32 else np.complex128)
        return a.reshape(-1, self.nlm), a.shape[:-1]

    def _grid_in(self, a, dtype):
        a = np.ascontiguousarray(a, dtype=np.float32 if dtype == DT_F32 else np.float64)
        return a.reshape(-1, self.nlat, self.nlon), a.shape[:-2]

    def spectogrd(self, spec, dtype=DT_F64):
        s, lead = self._spec_in(spec, dtype)
        g = np.zeros((s.shape[0], self.nlat, self.nlon), np.float32 if dtype else np.float64)
        _check(load().gfsdc_spectogrd(self.h, _p(s), _p(g), s.shape[0], dtype, HOST))
        return g.reshape(lead + g.shape[1:])

    def grdtospec(self, grid, dtype=DT_F64):
        g, lead = self._grid_in(grid, dtype)
        s = np.zeros((g.shape[0], self.nlm), np.complex64 if dtype else np.complex128)
        _check(load().gfsdc_grdtospec(self.h, _p(g), _p(s), g.shape[0], dtype, HOST))
        return s.reshape(lead + (self.nlm,))

    def getuv(self, vrt, div, rsphere, dtype=DT_F64):
        a, lead = self._spec_in(vrt, dtype)
        b, _ = self._spec_in(div, dtype)
        u = np.zeros((a.shape[0], self.nlat, self.nlon), np.float32 if dtype else np.float64)
        v = np.zeros_like(u)
        _check(load().gfsdc_getuv(self.h, _p(a), _p(b), _p(u), _p(v), a.shape[0], rsphere, dtype,
                                  HOST))
        return u.reshape(lead + u.shape[1:]), v.reshape(lead + v.shape[1:])

    def getvrtdivspec(self, u, v, rsphere, dtype=DT_F64):
        a, lead = self._grid_in(u, dtype)
        b, _ = self._grid_in(v, dtype)
        vr = np.zeros((a.shape[0], self.nlm), np.complex64 if dtype else np.complex128)
        dv = np.zeros_like(vr)
        _check(load().gfsdc_getvrtdivspec(self.h, _p(a), _p(b), _p(vr), _p(dv), a.shape[0],
                                          rsphere, dtype, HOST))
        return vr.reshape(lead + (self.nlm,)), dv.reshape(lead + (self.nlm,))

    def getgrad(self, spec, rsphere, dtype=DT_F64):
        a, lead = self._spec_in(spec, dtype)
        gx = np.zeros((a.shape[0], self.nlat, self.nlon), np.float32 if dtype else np.float64)
        gy = np.zeros_like(gx)
        _check(load().gfsdc_getgrad(self.h, _p(a), _p(gx), _p(gy), a.shape[0], rsphere, dtype,
                                    HOST))
        return gx.reshape(lead + gx.shape[1:]), gy.reshape(lead + gy.shape[1:])

    # -- dycore -------------------------------------------------------------
    def _s3(self, n=None):
        return np.zeros((self.nmem, n if n is not None else self.nlev, self.nlm), np.complex128)

    def set_state(self, vrt=None, div=None, tv=None, lnps=None, trac=None):
        arrs = [None if x is None else np.ascontiguousarray(x, np.complex128)
                for x in (vrt, div, tv, trac, lnps)]
        _check(load().gfsdc_set_state(self.h, *[_p(x) for x in arrs], DT_F64, HOST))

    def get_state(self):
        vrt, div, tv = self._s3(), self._s3(), self._s3()
        trac = np.zeros((self.nmem, max(self.ntrac, 1), self.nlev, self.nlm), np.complex128)
        lnps = np.zeros((self.nmem, self.nlm), np.complex128)
        _check(load().gfsdc_get_state(self.h, _p(vrt), _p(div), _p(tv),
                                      _p(trac) if self.ntrac else None, _p(lnps), DT_F64, HOST))
        return dict(vrt=vrt, div=div, tv=tv, trac=trac[:, :self.ntrac], lnps=lnps)

    def set_topo(self, topo):
        t = np.ascontiguousarray(np.broadcast_to(topo, (self.nmem, self.nlm)), np.complex128)
        _check(load().gfsdc_set_topo(self.h, _p(t), DT_F64, HOST))
        self._topo = t.copy()

    def topo(self):
        """spectral orography (nmem, nlm) as last set or restored"""
        t = getattr(self, "_topo", None)
        return np.zeros((self.nmem, self.nlm), np.complex128) if t is None else t.copy()

    @property
    def pdryini(self):
        out = np.zeros(self.nmem)
        _check(load().gfsdc_get_pdryini(self.h, _p(out)))
        return out

    @pdryini.setter
    def pdryini(self, v):
        a = np.ascontiguousarray(np.broadcast_to(np.asarray(v, np.float64), (self.nmem,)))
        _check(load().gfsdc_set_pdryini(self.h, _p(a)))

    def getdyntend(self, rkstage=1, inverse_only=False):
        out = dict(dvrt=self._s3(), ddiv=self._s3(), dtv=self._s3(),
                   dtrac=np.zeros((self.nmem, max(self.ntrac, 1), self.nlev, self.nlm),
                                  np.complex128),
                   dlnps=np.zeros((self.nmem, self.nlm), np.complex128))
        if inverse_only:
            _check(load().gfsdc_getdyntend(self.h, rkstage, 1, None, None, None, None, None,
                                           DT_F64, HOST))
            return None
        _check(load().gfsdc_getdyntend(self.h, rkstage, 0, _p(out["dvrt"]), _p(out["ddiv"]),
                                       _p(out["dtv"]), _p(out["dtrac"]) if self.ntrac else None,
                                       _p(out["dlnps"]), DT_F64, HOST))
        return out

    def advance(self, nsteps=1, use_graph=False):
        _check(load().gfsdc_advance(self.h, nsteps, int(use_graph)))

    def hs_phytend(self):
        a, b, c = self._s3(), self._s3(), self._s3()
        _check(load().gfsdc_hs_phytend(self.h, _p(a), _p(b), _p(c), DT_F64, HOST))
        return dict(dvrt=a, ddiv=b, dtv=c)

    def hs_drag_levels(self):
        """levels, counted from the lowest, on which the last lean Held-Suarez
        physics pass of advance formed the drag (0 before the first); above
        them p / ps < sigbot in every column, so the drag there is zero"""
        n = C.c_int(0)
        _check(load().gfsdc_hs_drag_levels(self.h, C.byref(n)))
        return n.value

    def simple_phytend(self):
        a, b, c = self._s3(), self._s3(), self._s3()
        q = np.zeros((self.nmem, max(self.ntrac, 1), self.nlev, self.nlm), np.complex128)
        _check(load().gfsdc_simple_phytend(self.h, _p(a), _p(b), _p(c), _p(q), DT_F64, HOST))
        return dict(dvrt=a, ddiv=b, dtv=c, dtrac=q[:, :self.ntrac])

    def pattern(self):
        """current AR(1) temperature pattern, (nmem, ndimspec) complex"""
        out = np.zeros((self.nmem, self.nlm), np.complex128)
        _check(load().gfsdc_get_pattern(self.h, _p(out), DT_F64, HOST))
        return out

    def autotune(self):
        """pick the FFT fields-per-workgroup variants by timing (synthesis and analysis
        separately); returns (synthesis choice, synthesis + analysis ms per variant 2/4/8)"""
        ms = np.zeros(3)
        best = C.c_int(0)
        _check(load().gfsdc_autotune(self.h, _p(ms), C.byref(best)))
        return best.value, ms

    def bench_transforms(self, nscalar, nvector, reps):
        """ms per spectogrd+grdtospec round trip of nscalar fields + getuv of nvector pairs"""
        ms = C.c_double(0)
        _check(load().gfsdc_bench_transforms(self.h, nscalar, nvector, reps, C.byref(ms)))
        return ms.value

    def bench_copy(self, nbytes, reps):
        """ms per device copy of nbytes (profiler counter calibration)"""
        ms = C.c_double(0)
        _check(load().gfsdc_bench_copy(self.h, nbytes, reps, C.byref(ms)))
        return ms.value

    # -- learned column physics ------------------------------------------------
    def mlp_init(self, hidden=512, nlayers=3, maxbatch=4096, capacity=1 << 20, seed=1):
        _check(load().gfsdc_mlp_init(self.h, hidden, nlayers, maxbatch, capacity, seed))
        self.mlp_nin, self.mlp_nout = 4 * self.nlev + 3, 4 * self.nlev

    def mlp_collect(self, nsteps, stride):
        n = C.c_long(0)
        _check(load().gfsdc_mlp_collect(self.h, nsteps, stride, C.byref(n)))
        return n.value

    def mlp_normalize(self):
        _check(load().gfsdc_mlp_normalize(self.h))

    def mlp_train(self, nsteps, batch, lr, wd=0.0, ntrain=0):
        losses = np.zeros(max(nsteps, 1))
        _check(load().gfsdc_mlp_train(self.h, nsteps, batch, lr, wd, ntrain, _p(losses)))
        return losses[:nsteps]

    def mlp_eval(self, first, count):
        e, b = C.c_double(0), C.c_double(0)
        _check(load().gfsdc_mlp_eval(self.h, first, count, C.byref(e), C.byref(b)))
        return e.value, b.value

    def mlp_enable(self, on=True, judge_stride=50, threshold=0.5):
        _check(load().gfsdc_mlp_enable(self.h, int(on), judge_stride, threshold))

    def mlp_judge_stats(self):
        out = np.zeros(3)
        _check(load().gfsdc_mlp_judge_stats(self.h, _p(out)))
        return dict(error=out[0], fallbacks=int(out[1]), judged=int(out[2]))

    def mlp_params(self, set=None):
        n = load().gfsdc_mlp_nparams(self.h)
        out = np.zeros(n, np.float32)
        src = None if set is None else np.ascontiguousarray(set, np.float32)
        _check(load().gfsdc_mlp_params(self.h, _p(out), _p(src) if src is not None else None))
        return out

    def mlp_norm(self, set=None):
        """the normalisation [mean_x | std_x | mean_y | std_y]; set= installs one
        (with mlp_params, a network trained in another context of the same levels)"""
        out = np.zeros(2 * (self.mlp_nin + self.mlp_nout), np.float32)
        src = None if set is None else np.ascontiguousarray(set, np.float32)
        if src is not None and src.size != out.size:
            raise GfsdcError(f"mlp_norm: {src.size} values, expected {out.size}")
        _check(load().gfsdc_mlp_norm(self.h, _p(out) if set is None else None, _p(src) if src is not None else None))
        return out if set is None else src

    def mlp_forward(self, x):
        x = np.ascontiguousarray(x, np.float32)
        y = np.zeros((x.shape[0], self.mlp_nout), np.float32)
        _check(load().gfsdc_mlp_forward(self.h, _p(x), x.shape[0], _p(y)))
        return y

    def mlp_grad(self, x, t):
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        g = np.zeros(load().gfsdc_mlp_nparams(self.h), np.float32)
        loss = C.c_double(0)
        _check(load().gfsdc_mlp_grad(self.h, _p(x), _p(t), x.shape[0], _p(g), C.byref(loss)))
        return g, loss.value

    def mlp_adamw(self, grad, lr, wd=0.0):
        g = None if grad is None else np.ascontiguousarray(grad, np.float32)
        _check(load().gfsdc_mlp_adamw(self.h, _p(g) if g is not None else None, lr, wd))

    def grid(self, name):
        nk = {"etadot": self.nlev + 1, "psg": 1, "dlnpsdt": 1, "dlnpsdx": 1, "dlnpsdy": 1,
              "tracerg": self.nlev * self.ntrac, "precip": 1, "pwat": 1}.get(name, self.nlev)
        out = np.zeros((self.nmem, nk, self.nlat, self.nlon))
        _check(load().gfsdc_get_grid(self.h, self.GRID[name], _p(out), DT_F64, HOST))
        return out

    def semimp(self):
        n = self.nlev
        A = np.zeros(n * n); B = np.zeros(n * n); sv = np.zeros(n); tor = np.zeros(n)
        D = np.zeros(n * n * (self.ntrunc + 1) * 3)
        _check(load().gfsdc_get_semimp(self.h, _p(A), _p(B), _p(sv), _p(tor), _p(D)))
        return dict(amhyb=A.reshape(n, n).T.copy(), bmhyb=B.reshape(n, n).T.copy(), svhyb=sv,
                    tor_hyb=tor,
                    d_hyb_m=D.reshape(3, self.ntrunc + 1, n, n).transpose(0, 1, 3, 2).copy())

    def damp(self):
        d = np.zeros(self.nlm); f = np.zeros(self.nlev); p = np.zeros(self.nlev)
        _check(load().gfsdc_get_damp(self.h, _p(d), _p(f), _p(p)))
        return dict(disspec=d, diff_prof=f, dmp_prof=p)

    def step_diag(self):
        out = np.zeros(4 * self.nmem)
        _check(load().gfsdc_step_diag(self.h, _p(out)))
        o = out.reshape(self.nmem, 4)
        return dict(spdmax=o[:, 0], psmin=o[:, 1], psmax=o[:, 2], pstend=o[:, 3])

    def init_rest_state(self, t0=300.0, ps0=1.0e5, vort_rms=1.0e-6, nmin=1, nmax=21, seed=0):
        _check(load().gfsdc_init_rest_state(self.h, t0, ps0, vort_rms, nmin, nmax, seed))

    @property
    def time(self):
        """model clock (s) at the start of the next step"""
        t = C.c_double(0.0)
        _check(load().gfsdc_get_time(self.h, C.byref(t)))
        return t.value

    @time.setter
    def time(self, t):
        _check(load().gfsdc_set_time(self.h, float(t)))

    def iau_set(self, fhrs, delthrs, vrt=None, div=None, tv=None, lnps=None, trac=None):
        """IAU increments (analysis - first guess) at the hours fhrs: vrt, div,
        tv (nfiles, nmem, nlev, nlm), trac (nfiles, nmem, ntrac, nlev, nlm),
        lnps (nfiles, nmem, nlm); fhrs empty switches the forcing off."""
        nf = len(fhrs)
        if nf == 0:
            _check(load().gfsdc_iau_set(self.h, 0, None, 0.0, None, None, None, None, None, DT_F64,
                                        HOST))
            return
        h = np.ascontiguousarray(fhrs, np.float64)
        a = [None if x is None else np.ascontiguousarray(x, np.complex128)
             for x in (vrt, div, tv, trac if self.ntrac else None, lnps)]
        _check(load().gfsdc_iau_set(self.h, nf, _p(h), float(delthrs), *[_p(x) for x in a], DT_F64,
                                    HOST))

    def dfi(self, ntdfi, use_graph=False):
        _check(load().gfsdc_dfi(self.h, ntdfi, int(use_graph)))

    STOCH = ["svc", "sppt", "shum", "svc_spec", "sppt_spec", "shum_spec", "addnoise", "psiforcing"]

    def stoch_init(self, **kw):
        """init_stochdata with the params.f90 stochastic settings given as keywords"""
        c = StochCfg()
        load().gfsdc_stoch_defaults(C.byref(c))
        for k, v in kw.items():
            setattr(c, k, v)
        self._stoch = c
        _check(load().gfsdc_stoch_init(self.h, C.byref(c)))

    def stoch(self, which):
        """'svc', 'sppt', 'shum' grids (nmem, nlat, nlon); their spectral patterns
        'svc_spec', 'sppt_spec', 'shum_spec' (nmem, nlm); 'addnoise' patterns and
        the last step's 'psiforcing' (nmem, nlev, nlm)"""
        w = self.STOCH.index(which)
        if w < 3:
            out = np.zeros((self.nmem, self.nlat, self.nlon))
        elif w < 6:
            out = np.zeros((self.nmem, self.nlm), np.complex128)
        else:
            out = np.zeros((self.nmem, self.nlev, self.nlm), np.complex128)
        _check(load().gfsdc_get_stoch(self.h, w, _p(out), DT_F64, HOST))
        return out

    def checkpoint(self):
        """the device state as one byte image (numpy uint8)"""
        n = C.c_size_t(0)
        _check(load().gfsdc_checkpoint_size(self.h, C.byref(n)))
        buf = np.zeros(n.value, np.uint8)
        _check(load().gfsdc_checkpoint_save(self.h, _p(buf), n.value, HOST))
        return buf

    def restore(self, buf):
        buf = np.ascontiguousarray(buf, np.uint8)
        _check(load().gfsdc_checkpoint_load(self.h, _p(buf), buf.size, HOST))
        # the orography section follows the 128-byte header and the spectral state
        c = np.complex128 if self.precision == FP64 else np.complex64
        nfs = self.nlev * (3 + self.ntrac) + 1
        off = 128 + np.dtype(c).itemsize * self.nlm * nfs * self.nmem
        n = self.nlm * self.nmem
        self._topo = buf[off:off + n * np.dtype(c).itemsize].view(c).astype(np.complex128).reshape(
            self.nmem, self.nlm)

    def read_sigma(self, path):
        """readin_sig and init_dyn (dyn_init.f90:43-57,120-149) with params.f90:
        241-264: the state, orography, pdryini and clock of every member from a
        sigma file written for this context's grid, truncation and levels."""
        import sigio
        head, data = sigio.read(path)
        if (head["jcap"], head["levs"]) != (self.ntrunc, self.nlev) or head["ntrac"] < self.ntrac:
            raise GfsdcError(f"{path}: T{head['jcap']}L{head['levs']} with {head['ntrac']} tracers, "
                             f"context is T{self.ntrunc}L{self.nlev} with {self.ntrac}")
        st = sigio.model_state(head, data, self)
        tol = 1e-6 * max(np.abs(self._ak).max(), 1.0)
        if np.abs(st["ak"] - self._ak).max() > tol or np.abs(st["bk"] - self._bk).max() > 1e-6:
            raise GfsdcError(f"{path}: vertical coordinate differs from the context's ak, bk")
        b = lambda x: np.ascontiguousarray(np.broadcast_to(x, (self.nmem,) + x.shape))
        self.set_state(b(st["vrt"]), b(st["div"]), b(st["tv"]), b(st["lnps"]),
                       b(st["trac"][:self.ntrac]) if self.ntrac else None)
        self.set_topo(st["topo"])
        self.pdryini = st["pdryini"]
        self.time = st["tstart"]
        return head

    def write_sigma(self, path, member=0, head=None):
        """wrtout_sig (dyn_init.f90:151-191): member's state at the current
        clock; head (a header as sigio.read returns) supplies the fields the
        model does not own, e.g. idate, else a modern hybrid header is made."""
        import sigio
        st = self.get_state()
        state = dict(vrt=st["vrt"][member], div=st["div"][member], tv=st["tv"][member],
                     trac=st["trac"][member], lnps=st["lnps"][member], topo=self.topo()[member])
        if head is None:
            head = sigio.header_for(self.nlon, self.nlat, self.ntrunc, self._ak, self._bk, self.ntrac)
        head = dict(head)
        head["fhour"] = self.time / 3600.0
        head["pdryini"] = float(self.pdryini[member]) / 1000.0
        sigio.write(path, head, sigio.sigma_data(state, self))

    def sync(self):
        _check(load().gfsdc_sync(self.h))

    def stream(self):
        return load().gfsdc_stream(self.h)

    def profile_step(self, nmax=64):
        names = (C.c_char * 48 * nmax)()
        ms = np.zeros(nmax)
        n = load().gfsdc_profile_step(self.h, names, _p(ms), nmax)
        if n < 0:
            _check(n)
        return [(bytes(names[k]).split(b"\0")[0].decode(), float(ms[k])) for k in range(n)]
