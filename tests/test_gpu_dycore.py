"""GPU parity of the dynamical core (libgfsdc) against the FP64 oracle.

Cases: T42L20 sigma (the CPU-runnable configuration) and a T42L64 hybrid
column set (the L64 coordinate of the larger configurations at a size the
oracle finishes in seconds). Tolerances are relative max-norm errors per
field: FP64 path 1e-9; FP32 path 2e-4 for tendencies (single-step
cancellation), and the multi-step FP32 drift is bounded in test_advance.
"""
import numpy as np
import pytest

import cases
import levels
import oracle as O

pytestmark = pytest.mark.gpu

NLON, NLAT, T = 128, 64, 42


def coords(kind, nlev):
    return levels.sigma_levels(nlev) if kind == "sigma" else levels.hybrid_levels(nlev)


def make_pair(kind, nlev, precision, dt=1200.0, **kw):
    import gfsdc
    ak, bk = coords(kind, nlev)
    o = O.Model(NLON, NLAT, T, nlev, ak, bk, dt, **kw)
    g = gfsdc.Context(NLON, NLAT, T, precision, nlev=nlev, ak=ak, bk=bk, dt=dt, **kw)
    return o, g


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("kind,nlev", [("sigma", 20), ("hybrid", 64)])
def test_operators_match(kind, nlev):
    import gfsdc
    o, g = make_pair(kind, nlev, gfsdc.FP64)
    so, sg = o.semimp(), g.semimp()
    for k in so:
        assert rel(sg[k], so[k]) < 1e-12, k
    do, dg = o.damp(), g.damp()
    for k in do:
        assert np.allclose(dg[k], do[k], rtol=1e-13, atol=0), k


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("kind,nlev", [("sigma", 20), ("hybrid", 64)])
def test_getdyntend(kind, nlev, prec):
    import gfsdc
    p = gfsdc.FP64 if prec == "fp64" else gfsdc.FP32
    o, g = make_pair(kind, nlev, p)
    vrt, div, tv, lnps = cases.perturbed_state(o.sht.degree, o.sht.order, nlev, seed=5)
    o.set_state(vrt[0], div[0], tv[0], lnps[0])
    g.set_state(vrt, div, tv, lnps)
    to = o.getdyntend(1)
    tg = g.getdyntend(1)
    tol = 1e-9 if prec == "fp64" else 2e-4
    for k in ("dvrt", "ddiv", "dtv", "dlnps"):
        assert rel(tg[k][0], to[k]) < tol, (k, rel(tg[k][0], to[k]))
    assert abs(g.pdryini[0] - o.pdryini) < (1e-9 if prec == "fp64" else 2e-6) * o.pdryini
    for name in ("ug", "vg", "virtempg", "dlnpdtg", "etadot", "psg", "dlnpsdt"):
        a, b = g.grid(name)[0], o.grid(name)
        assert rel(a, b) < (1e-9 if prec == "fp64" else 2e-4), (name, rel(a, b))


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_getdyntend_after_autotune(prec):
    """the tile autotuner times the FFT variants on the context's own buffers and
    leaves the model state alone: getdyntend still matches the oracle after it"""
    import gfsdc
    p = gfsdc.FP64 if prec == "fp64" else gfsdc.FP32
    o, g = make_pair("hybrid", 64, p)
    vrt, div, tv, lnps = cases.perturbed_state(o.sht.degree, o.sht.order, 64, seed=9)
    o.set_state(vrt[0], div[0], tv[0], lnps[0])
    g.set_state(vrt, div, tv, lnps)
    best, ms = g.autotune()
    assert best in (2, 4, 8) and np.all(ms > 0)
    to = o.getdyntend(1)
    tg = g.getdyntend(1)
    tol = 1e-9 if prec == "fp64" else 2e-4
    for k in ("dvrt", "ddiv", "dtv", "dlnps"):
        assert rel(tg[k][0], to[k]) < tol, (k, rel(tg[k][0], to[k]))


def test_rest_state_matches_numpy():
    import gfsdc
    o, g = make_pair("sigma", 20, gfsdc.FP64)
    g.init_rest_state(300.0, 1.0e5, 1e-6, 1, 21, seed=7)
    st = g.get_state()
    vrt, div, tv, lnps = cases.rest_state(o.sht.degree, o.sht.order, 20, seed=7)
    assert rel(st["vrt"], vrt) < 1e-14
    assert np.abs(st["div"]).max() == 0
    assert rel(st["tv"], tv) < 1e-15 and rel(st["lnps"], lnps) < 1e-15


@pytest.mark.parametrize("kind,nlev", [("sigma", 20), ("hybrid", 64)])
def test_advance_fp64(kind, nlev):
    import gfsdc
    o, g = make_pair(kind, nlev, gfsdc.FP64)
    vrt, div, tv, lnps = cases.perturbed_state(o.sht.degree, o.sht.order, nlev, seed=9, amp=0.3)
    o.set_state(vrt[0], div[0], tv[0], lnps[0])
    g.set_state(vrt, div, tv, lnps)
    o.advance(3)
    g.advance(3)
    so, sg = o.get_state(), g.get_state()
    for k in ("vrt", "div", "tv", "lnps"):
        assert rel(sg[k][0], so[k]) < 1e-9, (k, rel(sg[k][0], so[k]))
    assert abs(g.pdryini[0] - o.pdryini) < 1e-9 * o.pdryini
    dg, do = g.step_diag(), o.step_diag()
    for k in do:
        assert abs(dg[k][0] - do[k]) <= 1e-8 * abs(do[k]) + 1e-12, k


def test_advance_fp32_one_day_t42():
    """C0 recipe (T42L20 HS, dt 1200 s) for 72 steps: FP32 path vs FP64 oracle."""
    import gfsdc
    o, g = make_pair("sigma", 20, gfsdc.FP32)
    vrt, div, tv, lnps = cases.rest_state(o.sht.degree, o.sht.order, 20, seed=0)
    o.set_state(vrt[0], div[0], tv[0], lnps[0])
    g.init_rest_state(300.0, 1.0e5, 1e-6, 1, 21, seed=0)
    o.advance(72)
    g.advance(72)
    so, sg = o.get_state(), g.get_state()
    errs = {k: np.linalg.norm(sg[k][0] - so[k]) / np.linalg.norm(so[k])
            for k in ("vrt", "div", "tv", "lnps")}
    print("FP32 one-day spectral relative error (T42L20):", errs)
    # the accuracy bound of the FP32 path: <= 1e-5 per field after one model day
    assert all(e <= 1e-5 for e in errs.values()), errs


def test_advance_bf16x3_one_day_t42():
    """The same day with the BF16X3 Legendre GEMMs (three bf16 products per
    FP32 product, about 1e-5 per transform and up to 1e-3 for getuv): the state
    stays within 2e-4 of the FP64 oracle (measured 4.2e-5). A wiring error of
    the split GEMMs or of the Fourier-formed x-gradients would show as an O(1)
    difference."""
    import gfsdc
    o, g = make_pair("sigma", 20, gfsdc.BF16X3)
    vrt, div, tv, lnps = cases.rest_state(o.sht.degree, o.sht.order, 20, seed=0)
    o.set_state(vrt[0], div[0], tv[0], lnps[0])
    g.init_rest_state(300.0, 1.0e5, 1e-6, 1, 21, seed=0)
    o.advance(72)
    g.advance(72)
    so, sg = o.get_state(), g.get_state()
    errs = {k: np.linalg.norm(sg[k][0] - so[k]) / np.linalg.norm(so[k])
            for k in ("vrt", "div", "tv", "lnps")}
    print("BF16X3 one-day spectral relative error (T42L20):", errs)
    assert all(e <= 2e-4 for e in errs.values()), errs


def test_fp32_one_day_t126_hybrid_vs_fp64():
    """T126L64 on the hybrid coordinate, HS forcing, dt 600 s, one model day
    (144 steps): FP32 path against the FP64 GPU path (which equals the oracle to
    1e-9 per step in test_advance_fp64); <= 1e-5 per field."""
    import gfsdc
    ak, bk = levels.hybrid_levels(64)
    ctx = {}
    for p in (gfsdc.FP32, gfsdc.FP64):
        c = gfsdc.Context(384, 190, 126, p, nlev=64, ak=ak, bk=bk, dt=600.0)
        c.init_rest_state(300.0, 1.0e5, 1e-6, 1, 21, seed=0)
        c.advance(144, use_graph=True)
        ctx[p] = c.get_state()
    a, b = ctx[gfsdc.FP32], ctx[gfsdc.FP64]
    errs = {k: float(np.linalg.norm(a[k] - b[k]) / np.linalg.norm(b[k])) for k in ("vrt", "div", "tv", "lnps")}
    print("FP32 one-day spectral relative error (T126L64):", errs)
    assert all(e <= 1e-5 for e in errs.values()), errs


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_graph_replay_matches_eager(prec):
    import gfsdc
    p = gfsdc.FP64 if prec == "fp64" else gfsdc.FP32
    ak, bk = coords("hybrid", 64)
    mk = lambda: gfsdc.Context(NLON, NLAT, T, p, nlev=64, ak=ak, bk=bk, dt=600.0)
    a, b = mk(), mk()
    for c in (a, b):
        c.init_rest_state(300.0, 1.0e5, 1e-6, 1, 21, seed=3)
    a.advance(4, use_graph=False)
    b.advance(1, use_graph=True)
    b.sync()
    b.step_diag()  # eager work between replays
    b.advance(3, use_graph=True)
    sa, sb = a.get_state(), b.get_state()
    for k in ("vrt", "div", "tv", "lnps"):
        assert np.array_equal(sa[k], sb[k]), k


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_topography_enters_as_the_geopotential_gradient(prec):
    """set_topo (dyn_init.f90:84: dphisdx, dphisdy = getgrad(grav topo)) adds
    -grad(phi_s) to the pressure-gradient force and nothing else. With the same
    state, the tendencies with and without topography differ by exactly the
    divergence of that force, -lap(phi_s) = n(n+1)/a^2 phi_s, in ddiv; the
    vorticity, temperature and ln ps tendencies do not change (the curl of a
    gradient is zero). Ten members with distinct topographies put the y
    gradients of the last members into the next 16-field group of the
    Fourier buffer."""
    import gfsdc
    p = gfsdc.FP64 if prec == "fp64" else gfsdc.FP32
    nm = 10
    ak, bk = coords("hybrid", 20)
    ctx = [gfsdc.Context(NLON, NLAT, T, p, nlev=20, ak=ak, bk=bk, dt=1200.0, nmem=nm) for _ in range(2)]
    for c in ctx:
        c.init_rest_state(300.0, 1.0e5, 1e-6, 1, 21, seed=40)
    a = ctx[0]
    rng = np.random.default_rng(41)
    n = a.degree.astype(float)
    topo = 1.0e3 * (np.maximum(n, 1) ** -1.5)[None, :] * np.exp(2j * np.pi * rng.random((nm, a.nlm)))
    topo[:, a.order == 0] = topo[:, a.order == 0].real
    ctx[1].set_topo(topo)
    t0, t1 = ctx[0].getdyntend(1), ctx[1].getdyntend(1)
    want = (n * (n + 1) / a.cfg.rerth ** 2)[None, :] * (a.cfg.grav * topo)  # (nm, nlm)
    tol = 1e-9 if prec == "fp64" else 1e-4
    for mem in range(nm):
        scale = np.abs(want[mem]).max()
        dd = t1["ddiv"][mem] - t0["ddiv"][mem]
        assert np.abs(dd - want[mem][None, :]).max() < tol * scale, (mem, np.abs(dd - want[mem]).max() / scale)
        assert np.abs(t1["dvrt"][mem] - t0["dvrt"][mem]).max() < tol * scale, mem
        assert np.abs(t1["dtv"][mem] - t0["dtv"][mem]).max() == 0, mem
        assert np.abs(t1["dlnps"][mem] - t0["dlnps"][mem]).max() == 0, mem


@pytest.mark.parametrize("nfw", [2, 8])
def test_ensemble_members_independent(nfw, monkeypatch):
    """three members folded into every launch, each equal to its own
    single-member run; with 2 fields per Fourier workgroup (the general path)
    and 8 (the wavefront path), both of which map each member's x-gradients and
    temperature mean onto its own Legendre fields"""
    import gfsdc
    monkeypatch.setenv("GFSDC_FFT_NF", str(nfw))
    ak, bk = coords("sigma", 20)
    ens = gfsdc.Context(NLON, NLAT, T, gfsdc.FP64, nlev=20, ak=ak, bk=bk, dt=1200.0, nmem=3)
    ens.init_rest_state(300.0, 1.0e5, 1e-6, 1, 21, seed=100)
    ens.advance(2)
    se = ens.get_state()
    for mem in range(3):
        one = gfsdc.Context(NLON, NLAT, T, gfsdc.FP64, nlev=20, ak=ak, bk=bk, dt=1200.0)
        one.init_rest_state(300.0, 1.0e5, 1e-6, 1, 21, seed=100 + mem)
        one.advance(2)
        s1 = one.get_state()
        for k in ("vrt", "div", "tv", "lnps"):
            assert rel(se[k][mem], s1[k][0]) < 1e-12, (mem, k)
