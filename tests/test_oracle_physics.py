"""Simple moist physics of the oracle (oracle/phys.c, oracle/dyn.c
orc_simple_phytend) pinned three ways:
  * against the reference routine itself (simple_physics_v5_1.f90 compiled
    from its source into oracle/_ref by oracle/ref.mk; skipped where that
    library has not been built),
  * by conservation laws of the scheme (column water + precipitation,
    latent heating, flux-form boundary-layer diffusion),
  * by the grid wrapper's bookkeeping (phy_run_simple.f90:50-107) on a
    T42 model state.
"""
import ctypes as C
import os

import numpy as np
import pytest

import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
REF_SO = os.path.join(HERE, "..", "oracle", "_ref", "libref_simple_physics.so")

GRAVIT, RAIR, CPAIR, LATVAP, RH2O, RHOW = 9.80616, 287.0, 1.0045e3, 2.5e6, 461.5, 1000.0


def columns(ncol, nl, seed, moist=1.0, wind=30.0):
    """random but physically shaped columns, top -> bottom"""
    import levels
    rng = np.random.default_rng(seed)
    ak, bk = levels.hybrid_levels(nl)
    ps = 1.0e5 + rng.uniform(-4e3, 3e3, ncol)
    pint = ak[None, :] + bk[None, :] * ps[:, None]              # (ncol, nl+1)
    pint[:, 0] = np.maximum(pint[:, 0], 1.0)                    # finite top pressure
    pdel = np.diff(pint, axis=1)
    pmid = 0.5 * (pint[:, 1:] + pint[:, :-1])
    sig = pmid / ps[:, None]
    t = 200.0 + 100.0 * sig ** 0.3 + rng.normal(0, 2.0, sig.shape)
    es = 610.78 * np.exp(-LATVAP / RH2O * (1.0 / t - 1.0 / 273.16))
    qsat = RAIR / RH2O * es / pmid
    q = moist * qsat * rng.uniform(0.3, 1.3, sig.shape)         # some levels supersaturated
    u = rng.uniform(-wind, wind, sig.shape)
    v = rng.uniform(-wind, wind, sig.shape)
    lat = rng.uniform(-1.4, 1.4, ncol)
    return dict(t=t, q=q, u=u, v=v, pmid=pmid, pint=pint, pdel=pdel, rpdel=1.0 / pdel, ps=ps, lat=lat)


def run_oracle(c, dtime, sst_mode, turbflux):
    n = c["t"].shape[0]
    out = {k: np.zeros_like(c["t"]) for k in ("t", "q", "u", "v")}
    precl = np.zeros(n)
    for i in range(n):
        t, q, u, v, p = O.simple_column(dtime, c["lat"][i], c["t"][i], c["q"][i], c["u"][i], c["v"][i],
                                        c["pmid"][i], c["pint"][i], c["pdel"][i], c["rpdel"][i], c["ps"][i],
                                        sst_mode, turbflux)
        out["t"][i], out["q"][i], out["u"][i], out["v"][i], precl[i] = t, q, u, v, p
    return out, precl


def run_reference(c, dtime, test, turbflux):
    lib = C.CDLL(REF_SO)
    n, nl = c["t"].shape
    f = lambda a: np.asfortranarray(a, np.float64)  # (pcols, pver) column-major
    arr = {k: f(c[k]) for k in ("t", "q", "u", "v", "pmid", "pint", "pdel", "rpdel")}
    ps = np.array(c["ps"], np.float64)
    lat = np.array(c["lat"], np.float64)
    precl = np.zeros(n)
    ip = lambda x: C.byref(C.c_int(x))
    p = lambda a: a.ctypes.data_as(C.c_void_p)
    lib.simple_physics_(ip(n), ip(nl), C.byref(C.c_double(dtime)), p(lat), p(arr["t"]), p(arr["q"]),
                        p(arr["u"]), p(arr["v"]), p(arr["pmid"]), p(arr["pint"]), p(arr["pdel"]),
                        p(arr["rpdel"]), p(ps), p(precl), ip(test), ip(1 if turbflux else 0))
    return {k: np.ascontiguousarray(arr[k]) for k in ("t", "q", "u", "v")}, precl


@pytest.mark.skipif(not os.path.exists(REF_SO), reason="oracle/_ref not built (needs /root/reference)")
@pytest.mark.parametrize("nl,test,turbflux", [(64, 0, True), (64, 1, True), (20, 0, False), (30, 1, True)])
def test_column_matches_reference_routine(nl, test, turbflux):
    c = columns(48, nl, seed=nl + test)
    mine, pm = run_oracle(c, 600.0, test, turbflux)
    ref, pr = run_reference(c, 600.0, test, turbflux)
    for k in ("t", "q", "u", "v"):
        err = np.abs(mine[k] - ref[k]).max() / np.abs(ref[k]).max()
        assert err < 1e-13, (k, err)
    assert np.abs(pm - pr).max() <= 1e-13 * max(np.abs(pr).max(), 1e-30)
    assert (pr > 0).any()  # the random columns do condense


def test_condensation_conserves_water_and_energy():
    c = columns(32, 40, seed=3)
    dt = 900.0
    new, precl = run_oracle(c, dt, 0, False)
    w0 = (c["q"] * c["pdel"]).sum(1) / GRAVIT
    w1 = (new["q"] * c["pdel"]).sum(1) / GRAVIT
    # column water lost = precipitated water (kg/m^2): precl (m/s) * rho_w * dt
    np.testing.assert_allclose(w0 - w1, precl * RHOW * dt, rtol=1e-10, atol=1e-14)
    # latent heating: cp dT = L (-dq) level by level; winds untouched
    np.testing.assert_allclose(CPAIR * (new["t"] - c["t"]), -LATVAP * (new["q"] - c["q"]), rtol=1e-9, atol=1e-9)
    assert np.array_equal(new["u"], c["u"]) and np.array_equal(new["v"], c["v"])
    # condensing levels end close to saturation at their new temperature (a
    # linearised one-step adjustment, so within a few per cent)
    es = 610.78 * np.exp(-LATVAP / RH2O * (1.0 / new["t"] - 1.0 / 273.16))
    qs = RAIR / RH2O * es / c["pmid"]
    wet = new["q"] < c["q"]
    assert wet.any() and np.all(np.abs(new["q"][wet] / qs[wet] - 1.0) < 0.05)


def test_boundary_layer_diffusion_is_flux_form():
    """Dry columns (no condensation): after the implicit surface exchange on
    the lowest level, the tridiagonal diffusion conserves the mass-weighted
    column integrals of q and u exactly (CA(k) pdel(k) = CC(k+1) pdel(k+1))."""
    c = columns(24, 64, seed=5, moist=0.2)
    dt = 600.0
    new, precl = run_oracle(c, dt, 0, True)
    assert np.all(precl == 0.0)
    zvir = RH2O / RAIR - 1.0
    za = RAIR / GRAVIT * c["t"][:, -1] * (1 + zvir * c["q"][:, -1]) * 0.5 * (np.log(c["ps"]) - np.log(c["pint"][:, -2]))
    wind = np.hypot(c["u"][:, -1], c["v"][:, -1])
    cd = np.where(wind < 20.0, 0.0007 + 0.000065 * wind, 0.002)
    qsats = RAIR / RH2O * 610.78 / c["ps"] * np.exp(-LATVAP / RH2O * (1 / 302.15 - 1 / 273.16))
    q_s = c["q"].copy()
    q_s[:, -1] = (q_s[:, -1] + 0.0011 * wind * qsats * dt / za) / (1 + 0.0011 * wind * dt / za)
    u_s = c["u"].copy()
    u_s[:, -1] = u_s[:, -1] / (1 + cd * wind * dt / za)
    np.testing.assert_allclose((new["q"] * c["pdel"]).sum(1), (q_s * c["pdel"]).sum(1), rtol=1e-12)
    np.testing.assert_allclose((new["u"] * c["pdel"]).sum(1), (u_s * c["pdel"]).sum(1), rtol=1e-10,
                               atol=1e-9 * np.abs(c["u"] * c["pdel"]).sum(1).max())
    # diffusion smooths: the vertical variance of u inside the boundary layer does not grow
    bl = c["pmid"][0] > 85000.0
    assert np.var(new["u"][:, bl], axis=1).mean() <= np.var(u_s[:, bl], axis=1).mean()


def test_sst_modes():
    assert O.simple_sst(0, 0.3) == 302.15
    assert O.simple_sst(2, 0.0) == 302.15 and abs(O.simple_sst(2, np.pi / 2) - 272.15) < 1e-12
    # baroclinic-wave surface temperature: symmetric in latitude, warmest at the equator
    s = [O.simple_sst(1, x) for x in (-0.8, 0.0, 0.8)]
    assert abs(s[0] - s[2]) < 1e-12 and s[1] > s[0]


def test_grid_wrapper_bookkeeping():
    """phy_run_simple.f90:50-107 on a T42L20 moist state: tendencies are the
    column changes divided by dt, Tv <-> T with fv, pwat from dpk and q."""
    import cases
    import levels
    nl = 20
    ak, bk = levels.hybrid_levels(nl)
    m = O.Model(128, 64, 42, nl, ak, bk, 1200.0, ntrac=1, physics=2, simple_test=0, turbflux=1)
    vrt, div, tv, lnps = cases.perturbed_state(m.sht.degree, m.sht.order, nl, seed=2, amp=0.5)
    q = m.sht.grdtospec(cases.moist_grid(m.sht.lats, 128, ak, bk))
    m.set_state(vrt[0], div[0], tv[0], lnps[0], trac=q[None])
    m.getdyntend(3, inverse_only=True)
    out = m.simple_phytend()
    precip, pwat = m.grid("precip")[0], m.grid("pwat")[0]
    assert np.all(precip >= 0) and precip.max() > 0
    # pwat = sum(dpk * q_new) / g with q_new = q + dt * (grid tendency)
    dpk = m.grid("dpk")[::-1]                 # bottom -> top like the tracer
    qg = m.grid("tracerg")
    dq = m.sht.spectogrd(out["dtrac"][0]) * 1200.0
    pw = (dpk * (qg + dq)).sum(0) / m.cfg.grav
    # the analysis truncates the grid tendency: compare in a smoothed norm
    assert np.abs(pw - pwat).mean() < 0.05 * np.abs(pwat).mean()
    assert np.abs(out["dtrac"][0]).max() > 0 and np.abs(out["dtv"]).max() > 0
