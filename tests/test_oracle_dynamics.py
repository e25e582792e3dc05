"""Pins the oracle's dynamical core (restatement of src/dyn_run.f90,
src/pressure_data.f90, src/semimp_data.f90, src/run.f90, src/dyn_init.f90
setdampspec and src/phy_run_hs.f90) with closed forms and invariants that do
not come from the oracle:

* calc_pressdata in sigma coordinates against the closed forms
  pk = sigma ps, alfa = 1 - sigma_k/dsigma ln(sigma_k+1/sigma_k) and
  Phillips' layer pressure;
* the linearised getdyntend around a resting isothermal state against the
  independently assembled semi-implicit operators of init_semimpdata
  (amhyb, bmhyb, svhyb, tor_hyb); the one known mismatch, the top layer's
  ln(ps) coefficient alfa(1) = ln 2 (dyn_run.f90:393 with p_top = 0), is
  derived by hand and asserted;
* an exact steady state (solid-body zonal flow in gradient-wind balance,
  isothermal, sigma levels) whose tendencies must vanish (pins the Coriolis
  sign, the kinetic-energy term, the role swap of getvrtdivspec and the sign
  flip of the vorticity tendency);
* (I + c^2 dt^2 n(n+1) ym) d_hyb_m = I and the IMEX tableau identities;
* setdampspec against its closed form at GFS default efold;
* Held-Suarez forcing against the Held & Suarez (1994) formulas;
* a resting isothermal atmosphere stays at rest through advance.
The centred vertical-advection stencil (getvadv) has no closed form. It is
pinned by the constant-field checks and by its identity with the independently
coded flux-form scheme of getvadv_tracers on linear profiles, where the limiter
is inactive (every layer, top and bottom included).
"""
import numpy as np
import pytest

import cases
import levels
import oracle as O

NLON, NLAT, T = 128, 64, 42
SQ = np.sqrt(4 * np.pi)
F32LN2 = float(np.float32(np.log(2.0)))


def model(kind="sigma", nlev=20, dt=1200.0, **kw):
    ak, bk = levels.sigma_levels(nlev) if kind == "sigma" else levels.hybrid_levels(nlev)
    return O.Model(NLON, NLAT, T, nlev, ak, bk, dt, **kw), ak, bk


def rest(m, t0=300.0, ps=1e5):
    nl, nlm = m.nlev, m.nlm
    z = np.zeros((nl, nlm), np.complex128)
    tv = z.copy()
    tv[:, 0] = t0 * SQ
    lnps = np.zeros(nlm, np.complex128)
    lnps[0] = np.log(ps) * SQ
    return z.copy(), z.copy(), tv, lnps


def test_calc_pressdata_sigma_closed_form():
    nl = 20
    ak, bk = levels.sigma_levels(nl)
    kappa = 287.05 / 1004.6
    lnps = np.log(np.array([9.0e4, 1.0e5, 1.03e5]))
    p = O.calc_pressdata(ak, bk, kappa, lnps)
    ps = np.exp(lnps)
    sig = bk
    assert np.allclose(p["pk"], sig[:, None] * ps[None, :], rtol=1e-15, atol=0)
    assert np.allclose(p["dpk"], np.diff(sig)[:, None] * ps, rtol=1e-13)
    s1, s2 = sig[1:-1], sig[2:]
    alfa = 1 - s1 / (s2 - s1) * np.log(s2 / s1)
    assert np.allclose(p["alfa"][1:], alfa[:, None], rtol=1e-12)
    assert p["alfa"][0, 0] == F32LN2
    prs = ((sig[1:] ** (kappa + 1) - sig[:-1] ** (kappa + 1)) /
           ((kappa + 1) * np.diff(sig))) ** (1 / kappa)
    assert np.allclose(p["prs"][::-1], prs[:, None] * ps, rtol=1e-12)


def test_tableau_and_implicit_inverse():
    m, ak, bk = model("hybrid", 64, dt=600.0)
    s = m.semimp()
    A, B, sv, tor = s["amhyb"], s["bmhyb"], s["svhyb"], s["tor_hyb"]
    ym = np.outer(tor, sv) + A @ B
    c = [1 / 6, 1 / 3, 1 / 4]
    for st in range(3):
        for n in (0, 1, 17, 42):
            M = np.eye(64) + (float(np.float32(c[st])) * 600.0) ** 2 * n * (n + 1) * ym
            assert np.abs(M @ s["d_hyb_m"][st, n] - np.eye(64)).max() < 1e-10
    # semimp_data.f90:137-160 identities for the Lorenz-3 + trapezoidal tableau
    a21, a31, a32 = 1 / 3, 1 / 6, 1 / 2
    aa21, aa22, aa31, aa32, aa33 = 1 / 6, 1 / 6, 1 / 3, 0, 1 / 3
    b1, b2, b3, bb1, bb2, bb3, bb4 = 1 / 2, -1 / 2, 1, 3 / 8, 0, 3 / 8, 1 / 4
    assert abs(a21 - (aa21 + aa22)) < 1e-12 and abs(a31 + a32 - (aa31 + aa32 + aa33)) < 1e-12
    assert abs(b1 + b2 + b3 - 1) < 1e-12 and abs(bb1 + bb2 + bb3 + bb4 - 1) < 1e-12
    assert abs(a21 * b2 + (a31 + a32) * b3 - 0.5) < 1e-12
    assert abs((aa21 + aa22) * bb2 + (aa31 + aa32 + aa33) * bb3 + bb4 - 0.5) < 1e-12


def test_semimp_reference_column_closed_form():
    # sigma levels: svhyb = d sigma (bottom -> top), tor = R T / a^2, amhyb diagonal R alfa / a^2
    m, ak, bk = model("sigma", 20)
    s = m.semimp()
    c = m.cfg
    assert np.allclose(s["svhyb"], np.diff(bk)[::-1], rtol=1e-13)
    assert np.allclose(s["tor_hyb"], c.rd * 300.0 / c.rerth ** 2, rtol=1e-15)
    s1, s2 = bk[1:-1], bk[2:]
    alfa = np.concatenate([[F32LN2], 1 - s1 / (s2 - s1) * np.log(s2 / s1)])
    assert np.allclose(np.diag(s["amhyb"])[::-1], c.rd * alfa / c.rerth ** 2, rtol=1e-12)
    kappa = c.rd / c.cp
    assert np.allclose(np.diag(s["bmhyb"])[::-1], kappa * 300.0 * alfa, rtol=1e-12)


def test_rest_state_has_no_tendency():
    for kind, nl in (("sigma", 20), ("hybrid", 64)):
        m, _, _ = model(kind, nl)
        m.set_state(*rest(m))
        out = m.getdyntend(1)
        for k in ("dvrt", "ddiv", "dtv", "dlnps"):
            assert np.abs(out[k]).max() < 1e-20, (kind, k)


@pytest.mark.parametrize("kind,nl", [("sigma", 20), ("hybrid", 32)])
def test_linearisation_matches_semi_implicit_operators(kind, nl):
    m, ak, bk = model(kind, nl, omega=0.0)
    s = m.sht
    sm = m.semimp()
    A, B, sv, tor = sm["amhyb"], sm["bmhyb"], sm["svhyb"], sm["tor_hyb"]
    rng = np.random.default_rng(1)

    def rnd(scale):
        x = (rng.standard_normal(m.nlm) + 1j * rng.standard_normal(m.nlm)) * scale
        x[s.order == 0] = x[s.order == 0].real
        x[s.degree > 20] = 0
        x[0] = 0
        return x

    dT = np.array([rnd(1.0) for _ in range(nl)])
    dD = np.array([rnd(1e-6) for _ in range(nl)])
    dP = rnd(1e-3)
    for part in ("T", "D", "P"):
        res = []
        for eps in (1e-2, 1e-3):
            vrt, div, tv, lnps = rest(m, 300.0, 800e2)
            zT = eps * dT if part == "T" else 0 * dT
            zD = eps * dD if part == "D" else 0 * dD
            zP = eps * dP if part == "P" else 0 * dP
            m.set_state(vrt, div + zD, tv + zT, lnps + zP)
            m.pdryini = 1.0
            out = m.getdyntend(2)
            ddl = -s.lap[None, :] * (A @ zT + tor[:, None] * zP[None, :])
            if part == "P" and kind == "sigma":
                # top layer: the discrete pressure gradient carries alfa(1) = ln 2
                ddl[-1] *= F32LN2
            dtl = -(B @ zD)
            dpl = -(sv @ zD)
            scale = max(np.abs(ddl).max(), np.abs(dtl).max(), 1e-300)
            r = [np.abs(out["dtv"] - dtl).max() / max(np.abs(dtl).max(), 1e-300) if part == "D" else 0,
                 np.abs(out["dlnps"] - dpl).max() / max(np.abs(dpl).max(), 1e-300) if part == "D" else 0]
            if part != "D" and not (part == "P" and kind == "hybrid"):
                r.append(np.abs(out["ddiv"] - ddl).max() / scale)
            res.append(max(r))
        # residual is O(eps^2): shrinks ~10x per decade of eps (or sits at rounding level)
        assert res[1] < 1e-6 or res[1] < 0.2 * res[0], (part, res)
        assert res[1] < 1e-3, (part, res)


def test_balanced_zonal_flow_is_steady():
    m, ak, bk = model("sigma", 20)
    s, c = m.sht, m.cfg
    a, Om, R, T0, U = c.rerth, c.omega, c.rd, 300.0, 20.0
    mu = np.sin(s.lats)
    vrt, div, tv, _ = rest(m, T0)
    vrt[:, 1] = 2 * U / a / np.sqrt(3 / (4 * np.pi))
    lnpsg = np.log(1e5) - a * U * (2 * Om + U / a) / (2 * R * T0) * mu[:, None] ** 2 * np.ones((1, NLON))
    lnps = s.grdtospec(lnpsg)
    m.set_state(vrt, div, tv, lnps)
    out = m.getdyntend(1)
    scale = 2 * Om * U / a
    assert np.abs(out["dvrt"]).max() < 1e-8 * scale
    assert np.abs(out["ddiv"][:-1]).max() < 1e-8 * scale
    assert np.abs(out["dtv"][:-1]).max() < 1e-12
    assert np.abs(out["dlnps"]).max() < 1e-12
    top = -(F32LN2 - 1) * R * T0 * (s.lap / a ** 2) * lnps
    top[0] = 0
    assert np.abs(out["ddiv"][-1] - top).max() < 1e-8 * scale


def test_setdampspec_closed_form():
    m, ak, bk = model("sigma", 20)
    d = m.damp()
    s = m.sht
    a = m.cfg.rerth
    efold = 2.0 / (np.float32(3.e15) / a ** 4 * np.float32(80 * 81) ** 2)
    n = s.degree.astype(float)
    ref = -(1 / efold) * ((n * (n + 1)) / (42 * 43)) ** 4
    assert np.allclose(d["disspec"], ref, rtol=1e-6)
    assert np.allclose(d["diff_prof"], 1.0)  # fshk = 1 below T170
    assert np.all(d["dmp_prof"] == 0)        # sl > 0.002 everywhere


def test_held_suarez_forcing_closed_form():
    m, ak, bk = model("sigma", 20)
    s = m.sht
    vrt, div, tv, lnps = cases.perturbed_state(s.degree, s.order, 20, seed=2)
    m.set_state(vrt[0], div[0], tv[0], lnps[0])
    m.getdyntend(1, inverse_only=True)
    ref = m.hs_phytend()
    # Held & Suarez (1994): sigma_b = 0.7, k_f = 1/day, k_a = 1/40 day, k_s = 1/4 day,
    # dT_y = 60 K, dtheta_z = 10 K, T_eq >= 200 K, kappa = 2/7, p0 = 1000 hPa
    prs = m.grid("prs")
    ps = m.grid("psg")[0]
    sig = prs / ps
    lat = s.lats[None, :, None]
    kap = np.float32(2.0 / 7.0)
    teq = np.maximum(200.0, (315 - 60 * np.sin(lat) ** 2 - 10 * np.log(prs / 1e5) * np.cos(lat) ** 2) *
                     (prs / 1e5) ** kap)
    bl = np.maximum(0.0, (sig - np.float32(0.7)) / (1 - np.float32(0.7)))
    ka, ks, kf = np.float32(1 / (40 * 86400.)), np.float32(1 / (4 * 86400.)), np.float32(1 / 86400.)
    kt = ka + (ks - ka) * bl * np.cos(lat) ** 4
    ft = kt * (teq - m.grid("virtempg"))
    fv = -kf * bl * m.grid("vrtg")
    fd = -kf * bl * m.grid("divg")
    for k, g in (("dtv", ft), ("dvrt", fv), ("ddiv", fd)):
        want = s.grdtospec(g)
        assert np.abs(ref[k] - want).max() <= 1e-7 * np.abs(want).max(), k


def test_resting_atmosphere_stays_at_rest():
    m, ak, bk = model("hybrid", 20, dt=1200.0, physics=0)
    m.set_state(*rest(m))
    m.advance(3)
    st = m.get_state()
    assert np.abs(st["vrt"]).max() < 1e-25 and np.abs(st["div"]).max() < 1e-25
    assert np.abs(st["tv"][:, 1:]).max() < 1e-12 and np.abs(st["lnps"][1:]).max() < 1e-14


def test_vertical_advection_constant_and_linear():
    # u = U(level) with no divergence: etadot = 0 -> no vertical advection; a uniform
    # field has no advection whatever etadot is.
    m, ak, bk = model("sigma", 20, omega=0.0)
    s = m.sht
    vrt, div, tv, lnps = rest(m)
    div[:, 1] = 1e-6  # etadot != 0
    m.set_state(vrt, div, tv, lnps)
    out = m.getdyntend(1)
    # isothermal: dT/dt = kappa T omega/p only (no horizontal or vertical T advection)
    kappa = m.cfg.rd / m.cfg.cp
    want = s.grdtospec(kappa * m.grid("dlnpdtg") * 300.0)
    assert np.abs(out["dtv"] - want).max() < 1e-9 * np.abs(want).max()


def test_mass_fixer_restores_dry_mass():
    m, ak, bk = model("sigma", 20, dt=1200.0)
    vrt, div, tv, lnps = cases.rest_state(m.sht.degree, m.sht.order, 20, seed=0)
    m.set_state(vrt[0], div[0], tv[0], lnps[0])
    m.advance(2)
    p0 = m.pdryini
    m.getdyntend(1, inverse_only=True)
    ps = m.grid("psg")[0]
    assert abs((ps * m.sht.areawts).sum() - p0) < 1e-9 * p0


def _vadv_columns(seed, nl=30, ncol=400):
    rng = np.random.default_rng(seed)
    dpk = rng.uniform(500.0, 4000.0, (nl, ncol))
    et = np.zeros((nl + 1, ncol))
    et[1:-1] = rng.normal(0.0, 0.5, (nl - 1, ncol))   # Pa/s; zero at top and surface
    return dpk, et


def test_vadv_tracers_invariants():
    """getvadv_tracers (dyn_run.f90:478-603): the flux part telescopes (etadot
    vanishes at both ends), a constant field is not advected, and on a profile
    linear in the level index the limiter is inactive in the interior, where
    the scheme equals the centred getvadv difference (dyn_run.f90:453-476)."""
    nl = 30
    dpk, et = _vadv_columns(1, nl)
    rng = np.random.default_rng(2)
    q = np.abs(rng.normal(1e-3, 5e-4, (nl, dpk.shape[1])))
    v = O.vadv_tracers(q, et, dpk)
    # sum_K dpk_K vadv_K = sum_K D_K (et_K - et_{K+1})  (flux form plus continuity part)
    D = q[::-1]                       # top -> bottom
    lhs = (dpk * v[::-1]).sum(0)
    rhs = (D * (et[:-1] - et[1:])).sum(0)
    np.testing.assert_allclose(lhs, rhs, rtol=1e-10, atol=1e-14)
    assert np.abs(O.vadv_tracers(np.full_like(q, 3.7e-3), et, dpk)).max() < 1e-15
    lin = (1e-3 + 1e-5 * np.arange(nl))[:, None] * np.ones((1, dpk.shape[1]))
    v = O.vadv_tracers(lin, et, dpk)[::-1]            # top -> bottom
    Dl = lin[::-1]
    centred = 0.5 / dpk[1:-1] * (et[2:-1] * (Dl[2:] - Dl[1:-1]) + et[1:-2] * (Dl[1:-1] - Dl[:-2]))
    np.testing.assert_allclose(v[1:-1], centred, rtol=1e-9, atol=1e-18)


def test_vadv_equals_the_flux_scheme_on_linear_profiles():
    """getvadv (dyn_run.f90:453-476, the centred advective difference) against
    the independently coded flux-form scheme of getvadv_tracers (dyn_run.f90:
    478-603): on a profile linear in the level index whose values and end
    extrapolations keep one sign, every limiter ratio is 1, the limited
    interface values are the midpoints, and the flux form with its continuity
    term reduces to the centred difference in every layer, the top and bottom
    ones included, for either sign of etadot."""
    nl = 30
    dpk, et = _vadv_columns(5, nl)
    ncol = dpk.shape[1]
    rng = np.random.default_rng(6)
    a = rng.uniform(1.0, 2.0, ncol)
    b = rng.uniform(-0.02, 0.02, ncol)
    lin = a[None, :] + b[None, :] * np.arange(nl)[:, None]       # bottom -> top, positive
    np.testing.assert_allclose(O.vadv(lin, et, dpk), O.vadv_tracers(lin, et, dpk), rtol=1e-12, atol=1e-15)
    # a uniform field is not advected, whatever etadot is
    assert np.abs(O.vadv(np.full((nl, ncol), 3.0), et, dpk)).max() == 0
    # and a non-linear profile is where the two schemes part (the limiter acts)
    wavy = 1.5 + 0.4 * np.sin(0.9 * np.arange(nl))[:, None] * np.ones((1, ncol))
    assert np.abs(O.vadv(wavy, et, dpk) - O.vadv_tracers(wavy, et, dpk)).max() > 1e-6


def test_vadv_tracers_keeps_tracer_positive():
    """Thuburn's limiter is positive definite: a forward step with the
    vertical advection alone keeps a non-negative tracer (with zeros in it)
    non-negative for Courant numbers below 1/2."""
    nl = 40
    dpk, et = _vadv_columns(3, nl, ncol=300)
    rng = np.random.default_rng(4)
    q = np.where(rng.random((nl, dpk.shape[1])) < 0.3, 0.0, rng.uniform(0, 1e-2, (nl, dpk.shape[1])))
    v = O.vadv_tracers(q, et, dpk)
    cmax = (np.maximum(np.abs(et[:-1]), np.abs(et[1:])) / dpk).max()
    dt = 0.45 / cmax
    qn = q - dt * v
    assert qn.min() > -1e-12 * q.max()
    # an unlimited centred flux on the same columns does go negative
    D = q[::-1]
    half = 0.5 * (D[1:] + D[:-1])
    fl = np.zeros((nl + 1, q.shape[1]))
    fl[1:-1] = half * et[1:-1]
    vc = ((fl[1:] - fl[:-1]) + D * (et[:-1] - et[1:])) / dpk
    assert (D - dt * vc).min() < -1e-6 * q.max()


def test_topography_enters_as_the_geopotential_gradient():
    """dyn_init.f90:84 (dphisdx, dphisdy = getgrad(grav topo) / a) feeds
    -grad(phi_s) into the pressure-gradient force of getpresgrad and nothing
    else: with the same state the divergence tendency changes by exactly
    -lap(phi_s) = n(n+1)/a^2 g topo, the vorticity tendency by rounding (the
    curl of a gradient), temperature and ln ps not at all."""
    ak, bk = levels.hybrid_levels(12)
    ms = [O.Model(64, 32, 21, 12, ak, bk, 1200.0) for _ in range(2)]
    s = ms[0].sht
    vrt, div, tv, lnps = cases.rest_state(s.degree, s.order, 12, seed=3, vort_rms=1e-5)
    rng = np.random.default_rng(12)
    n = s.degree.astype(float)
    topo = 1.0e3 * np.maximum(n, 1) ** -1.5 * np.exp(2j * np.pi * rng.random(s.nlm))
    topo[s.order == 0] = topo[s.order == 0].real
    topo[0] = 0.0
    ms[1].set_topo(topo)
    for m in ms:
        m.set_state(vrt[0], div[0], tv[0], lnps[0])
    t0, t1 = ms[0].getdyntend(1), ms[1].getdyntend(1)
    c = ms[0].cfg
    want = n * (n + 1) / c.rerth ** 2 * (c.grav * topo)
    scale = np.abs(want).max()
    assert np.abs(t1["ddiv"] - t0["ddiv"] - want[None, :]).max() < 1e-12 * scale
    assert np.abs(t1["dvrt"] - t0["dvrt"]).max() < 1e-12 * scale
    assert np.abs(t1["dtv"] - t0["dtv"]).max() == 0 and np.abs(t1["dlnps"] - t0["dlnps"]).max() == 0
