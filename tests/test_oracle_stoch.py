"""Stochastic physics in the oracle (stoch_data.f90, patterngenerator.f90,
run.f90:272-292,509-553, dyn_run.f90:84-117,230-248), pinned by:

1. Patterns: the vertical factors against their closed forms (the sigma
   0.10-0.05 taper; exp((sl - 1) / 0.2) for shum); the sppt grid is the logit
   transform 2 / (1 + e^g) - 1 = -tanh(g / 2) of the synthesised pattern; each
   kind draws its own stream (same seed, different patterns) and a kind is
   reproducible from its seed; at stationarity the lag-1 autocorrelation of
   the coefficients is phi = exp(-dt / tau) and the grid variance is stdev^2
   (statistical, over many coefficients and steps).
2. Vorticity confinement: off whenever vcamp = 0, svc or not; with vcamp the
   tendency change is the vorticity / divergence of the flux
   F = vc |vrt| grad(vrt) / |grad(vrt)| (into ddiv / -dvrt, the reference's
   swapped getvrtdivspec), built here in numpy from the transforms; the
   confinement produces enstrophy, Int vrt d(vrt)/dt = Int vc |vrt| |grad vrt| > 0;
   svc multiplies vc by 1 + vfact grd_svc.
3. Additive noise: the 1-2-1 vertical filter keeps a level-constant pattern
   and maps the 2-level (+,-) pattern to 0 inside and to 1/3 of it at the
   ends (factor 4/3 there); without addnoise_dissfact the forcing is the
   filtered pattern times vfact; with it the forcing is homogeneous of degree
   1 in the wind (sqrt of the product of the wind and the dissipation wind)
   and vanishes at rest; the step adds dt lap / a^2 psiforcing to vorticity
   (dt sqrt(-lap) / a with addnoise_kenorm) and changes nothing else.
"""
import numpy as np
import pytest

import cases
import levels
import oracle as O

NLON, NLAT, T, NL = 64, 32, 21, 12
PAT = dict(tau=3 * 3600.0, lscale=5.0e5)


def model(physics=1, dt=1200.0, **stoch):
    ak, bk = levels.hybrid_levels(NL)
    m = O.Model(NLON, NLAT, T, NL, ak, bk, dt, physics=physics)
    if stoch:
        m.stoch_init(**stoch)
    return m


def kind(name, amp, seed=7, tau=PAT["tau"], lscale=PAT["lscale"]):
    return {name: amp, name + "_tau": tau, name + "_lscale": lscale, "iseed_" + name: seed}


def state(m, seed=2, vort_rms=2e-5):
    s = m.sht
    vrt, div, tv, lnps = cases.rest_state(s.degree, s.order, NL, seed=seed, vort_rms=vort_rms, nmax=15)
    return vrt[0], div[0], tv[0], lnps[0]


def sl_of(m):
    ak, bk = levels.hybrid_levels(NL)
    si = (ak / np.float32(101300.0) + bk)[::-1]
    return 0.5 * (si[:-1] + si[1:])


def test_vertical_factors_closed_form():
    m = model(**kind("svc", 0.5), **kind("shum", 0.01), **kind("sppt", 0.3), **kind("addnoise", 1e5))
    slv = sl_of(m)
    taper = np.where(slv <= 0.05, 0.0, np.where(slv < 0.10, (slv - 0.05) / 0.05, 1.0))
    for k in ("svc", "sppt", "addnoise"):
        np.testing.assert_allclose(m.stoch(k + "_vfact"), taper, rtol=1e-5, atol=1e-7)
    np.testing.assert_allclose(m.stoch("shum_vfact"), np.exp((slv - 1.0) / 0.2), rtol=1e-5)


def test_sppt_grid_is_the_logit_of_the_pattern_and_streams_differ():
    m = model(**kind("sppt", 0.8, seed=3), **kind("svc", 0.8, seed=3), **kind("shum", 0.8, seed=3))
    m.set_state(*state(m))
    m.advance(1)
    g = m.sht.spectogrd(m.stoch("sppt_spec")[None])[0]
    np.testing.assert_allclose(m.stoch("sppt"), -np.tanh(g / 2), rtol=0, atol=1e-13)
    np.testing.assert_allclose(m.stoch("svc"), m.sht.spectogrd(m.stoch("svc_spec")[None])[0], atol=1e-13)
    a, b, c = m.stoch("sppt_spec"), m.stoch("svc_spec"), m.stoch("shum_spec")
    assert np.abs(a - b).max() > 0.1 * np.abs(a).max() and np.abs(b - c).max() > 0.1 * np.abs(b).max()
    # a kind is reproducible from its seed, whatever else is on
    m2 = model(**kind("sppt", 0.8, seed=3))
    m2.set_state(*state(m2))
    m2.advance(1)
    np.testing.assert_array_equal(m2.stoch("sppt_spec"), a)


def test_pattern_autocorrelation_and_variance():
    dt, tau, amp = 1200.0, 2 * 3600.0, 0.7
    m = model(physics=0, dt=dt, **kind("svc", amp, tau=tau, lscale=3.0e5))
    m.set_state(*state(m, vort_rms=1e-7))
    phi = np.exp(-dt / tau)
    prev = m.stoch("svc_spec")
    num = den = 0.0
    var = []
    for _ in range(60):
        m.advance(1)
        cur = m.stoch("svc_spec")
        num += np.real(np.vdot(prev[1:], cur[1:]))
        den += np.real(np.vdot(prev[1:], prev[1:]))
        g = m.stoch("svc")
        var.append((m.sht.areawts.reshape(NLAT, NLON) * g ** 2).sum())
        prev = cur
    assert abs(num / den - phi) < 0.03, (num / den, phi)
    assert 0.75 * amp ** 2 < np.mean(var) < 1.25 * amp ** 2, (np.mean(var), amp ** 2)


def _flux_change(m, vc_grid):
    """(Delta ddiv, Delta dvrt) of the confinement flux on the model's state, in numpy"""
    s = m.sht
    st = m.get_state()
    a = m.cfg.rerth
    gx, gy = s.getgrad(st["vrt"], a)
    z = s.spectogrd(st["vrt"])
    ag = np.sqrt(gx ** 2 + gy ** 2)
    tiny = np.finfo(np.float32).tiny
    w = np.where(ag > tiny, 1.0 / np.where(ag > tiny, ag, 1.0), 0.0)
    fx = vc_grid * gx * w * np.abs(z)
    fy = vc_grid * gy * w * np.abs(z)
    vf, df = s.getvrtdivspec(fx, fy, a)
    return vf, -df


def test_vorticity_confinement_flux():
    m0 = model()
    y = state(m0)
    m0.set_state(*y)
    t0 = m0.getdyntend(1)
    # svc alone does nothing: it modulates vcamp = 0
    ms = model(**kind("svc", 0.5))
    ms.set_state(*y)
    ts = ms.getdyntend(1)
    for k in t0:
        np.testing.assert_array_equal(ts[k], t0[k])
    vcamp = 8.0
    m = model(vcamp=vcamp)
    m.set_state(*y)
    t1 = m.getdyntend(1)
    dd, dv = _flux_change(m, vcamp)
    sc = np.abs(dd).max()
    assert np.abs(t1["ddiv"] - t0["ddiv"] - dd).max() < 1e-11 * sc
    assert np.abs(t1["dvrt"] - t0["dvrt"] - dv).max() < 1e-11 * np.abs(dv).max()
    np.testing.assert_array_equal(t1["dtv"], t0["dtv"])
    # enstrophy production: sum over coefficients of vrt* d(vrt)/dt (orthonormal, m > 0 twice)
    s = m.sht
    wgt = np.where(s.order == 0, 1.0, 2.0)[None, :]
    prod = np.real(wgt * np.conj(y[0]) * (t1["dvrt"] - t0["dvrt"])).sum()
    gx, gy = s.getgrad(y[0], m.cfg.rerth)
    z = s.spectogrd(y[0])
    want = 4 * np.pi * (s.areawts.reshape(NLAT, NLON)[None] * vcamp * np.abs(z) * np.sqrt(gx ** 2 + gy ** 2)).sum()
    assert prod > 0 and 0.8 < prod / want < 1.2, (prod, want)


def test_stochastic_vorticity_confinement_modulation():
    vcamp = 6.0
    m = model(vcamp=vcamp, **kind("svc", 0.8, seed=11))
    y = state(m)
    m.set_state(*y)
    m0 = model()
    m0.set_state(*y)
    t0 = m0.getdyntend(1)
    t1 = m.getdyntend(1)
    vc = vcamp * (1.0 + m.stoch("svc_vfact")[:, None, None] * m.stoch("svc")[None])
    dd, dv = _flux_change(m, vc)
    assert np.abs(t1["ddiv"] - t0["ddiv"] - dd).max() < 1e-11 * np.abs(dd).max()
    # and the modulation matters: the pattern is not near zero
    dd0, _ = _flux_change(m, vcamp)
    assert np.abs(dd - dd0).max() > 0.1 * np.abs(dd0).max()


def test_addnoise_vertical_filter():
    m = model(**kind("addnoise", 1e5), addnoise_vfilt=1)
    m.set_state(*state(m))
    m.getdyntend(3, inverse_only=True)
    s = m.sht
    rng = np.random.default_rng(1)
    q = (rng.standard_normal(s.nlm) + 1j * rng.standard_normal(s.nlm)) * (s.degree > 0) * (s.degree < 10)
    q[s.order == 0] = q[s.order == 0].real
    vf = m.stoch("addnoise_vfact")
    full = vf == 1.0
    const = np.tile(q, (NL, 1))
    m.set_stoch("addnoise", const)
    out = m.stoch_forcing()
    # interior rows exactly; the end rows to the float32 rounding of the literal 1 + 1/3
    inner = full.copy()
    inner[[0, -1]] = False
    np.testing.assert_allclose(out[inner], const[inner], rtol=0, atol=1e-12 * np.abs(q).max())
    np.testing.assert_allclose(out[full], const[full], rtol=0, atol=1e-7 * np.abs(q).max())
    alt = const * ((-1.0) ** np.arange(NL))[:, None]
    m.set_stoch("addnoise", alt)
    out = m.stoch_forcing()
    assert np.abs(out[inner]).max() < 1e-12 * np.abs(q).max()
    np.testing.assert_allclose(out[0], alt[0] / 3.0 * vf[0], atol=1e-7 * np.abs(q).max())
    # without the filter and without dissfact: the pattern times vfact
    m2 = model(**kind("addnoise", 1e5))
    m2.set_state(*state(m2))
    m2.getdyntend(3, inverse_only=True)
    m2.set_stoch("addnoise", alt)
    np.testing.assert_allclose(m2.stoch_forcing(), alt * vf[:, None], atol=1e-12 * np.abs(q).max())


def test_addnoise_dissipation_modulation():
    m = model(**kind("addnoise", 1e5), addnoise_dissfact=1)
    s = m.sht
    rng = np.random.default_rng(3)
    pat = (rng.standard_normal((NL, s.nlm)) + 1j * rng.standard_normal((NL, s.nlm))) * (s.degree > 0)
    pat[:, s.order == 0] = pat[:, s.order == 0].real
    y = state(m, vort_rms=3e-5)
    out = []
    for f in (1.0, 2.0):
        m.set_state(y[0] * f, y[1] * f, y[2], y[3])
        m.getdyntend(3, inverse_only=True)
        m.set_stoch("addnoise", pat)
        out.append(m.stoch_forcing())
    assert np.abs(out[1] - 2.0 * out[0]).max() < 1e-9 * np.abs(out[1]).max()
    # at rest the dissipation rate is zero and so is the forcing (up to sqrt(tiny))
    m.set_state(0 * y[0], 0 * y[1], y[2], y[3])
    m.getdyntend(3, inverse_only=True)
    m.set_stoch("addnoise", pat)
    assert np.abs(m.stoch_forcing()).max() < 1e-15 * np.abs(out[0]).max()


@pytest.mark.parametrize("kenorm", [0, 1])
def test_addnoise_enters_vorticity_only(kenorm):
    dt = 900.0
    kw = dict(**kind("addnoise", 2e5, seed=5), addnoise_vfilt=1, addnoise_dissfact=1, addnoise_kenorm=kenorm)
    m, m0 = model(dt=dt, **kw), model(dt=dt)
    y = state(m)
    m.set_state(*y)
    m0.set_state(*y)
    m.advance(1)
    m0.advance(1)
    a, b = m.get_state(), m0.get_state()
    psi = m.stoch("psiforcing")
    lap, rs = m.sht.lap, m.cfg.rerth
    want = -dt * np.sqrt(-lap) / rs * psi if kenorm else dt * lap / rs ** 2 * psi
    assert np.abs(want).max() > 1e-3 * np.abs(b["vrt"]).max()
    scale = max(np.abs(want).max(), np.abs(b["vrt"]).max())
    assert np.abs(a["vrt"] - b["vrt"] - want).max() < 1e-12 * scale
    for k in ("div", "tv", "lnps"):
        np.testing.assert_array_equal(a[k], b[k])
