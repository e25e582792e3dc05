"""Pins the oracle's getomega (dyn_run.f90:307-373) and getpresgrad
(dyn_run.f90:375-451) on general columns: non-isothermal temperature,
level-dependent winds, non-zero V.grad ln ps, sigma and hybrid levels.

Everything the checks compare against is built here in numpy from the
Simmons & Burridge (1981) hybrid-coordinate discretisation stated in its
physical form, not from the reference's cofa/cofb/workb/workc recurrences:

* interface pressures p_K = A_K + B_K ps (top -> bottom), layer thicknesses
  dp_K, ln(p_K+1 / p_K), alfa_K = 1 - p_K/dp_K ln(p_K+1/p_K), alfa_1 = ln 2;
* the hydrostatic geopotential at full levels,
  phi_K = phi_s + sum_{J below K} R T_J ln(p_J+1/p_J) + alfa_K R T_K;
* every horizontal gradient of a pressure-dependent quantity by complex-step
  differentiation of that quantity along (ps d ln ps/dx, dT/dx), which is
  exact to rounding and shares no algebra with the reference's closed forms.

Checks fixed by the mathematics:

* continuity: ps d ln ps/dt = -sum_K div(V_K dp_K), and each layer's
  discrete mass budget d dp_K/dt + div(V_K dp_K) + etadot_K+1 - etadot_K = 0;
* the S&B energy-conversion identity tying omega/p to the hydrostatic
  geopotential: sum_K R T_K dp_K (V_K.(grad ln p)_K - (omega/p)_K)
  = sum_K (phi_K - phi_s) div(V_K dp_K);
* the pressure-gradient force -(grad phi_K + R T_K (grad ln p)_K) per level;
  in sigma coordinates (grad ln p)_K = grad ln ps exactly (the textbook form);
* exact balance over an isothermal column: the force is -grad phi_s -
  R T0 grad ln ps on every level below the top, for any hybrid levels;
* the column angular-momentum identity of S&B:
  sum_K dp_K PGF_K = -grad(sum_K dp_K phi_K) + phi_s grad ps.

The reference's top layer (p_top = 0) keeps hand-derived special forms,
asserted as such: alfa_1 = ln 2 (float literal, pressure_data.f90:92) in the
geopotential and pressure gradient; omega/p's advective part V.grad(dp_1)/dp_1
without alfa_1 (dyn_run.f90:358); and alfa_1 multiplying div_1 dp_1 +
ps V_1.grad ln ps dB_1^2 (dyn_run.f90:349-350 use cb(1)*dbk(1), where the lower
layers use cg(k)*dbk(k)), which differs from S&B on sigma levels only.
"""
import numpy as np
import pytest

import levels
import oracle as O

RD = 287.05
F32LN2 = float(np.float32(np.log(2.0)))
H = 1e-30  # complex-step size


def columns(kind, nl, ncol=300, seed=0):
    ak, bk = levels.sigma_levels(nl) if kind == "sigma" else levels.hybrid_levels(nl)
    rng = np.random.default_rng(seed)
    lnps = np.log(rng.uniform(6.0e4, 1.04e5, ncol))
    # non-isothermal, level- and column-dependent temperature and winds (bottom -> top)
    z = np.linspace(0.0, 1.0, nl)[:, None]
    tv = 290.0 - 80.0 * z + rng.normal(0.0, 6.0, (nl, ncol))
    ug = rng.normal(0.0, 20.0, (nl, ncol)) + 30.0 * z
    vg = rng.normal(0.0, 15.0, (nl, ncol))
    divg = rng.normal(0.0, 1e-5, (nl, ncol))
    d = dict(
        dlnpsdx=rng.normal(0.0, 2e-6, ncol), dlnpsdy=rng.normal(0.0, 2e-6, ncol),
        dtvdx=rng.normal(0.0, 2e-5, (nl, ncol)), dtvdy=rng.normal(0.0, 2e-5, (nl, ncol)),
        phis=rng.uniform(0.0, 3.0e4, ncol),
        dphisdx=rng.normal(0.0, 5e-3, ncol), dphisdy=rng.normal(0.0, 5e-3, ncol))
    return ak, bk, lnps, tv, ug, vg, divg, d


# ---- the S&B discretisation in its physical form (top -> bottom, K = 0..nl-1) ----

def p_int(ak, bk, ps):
    return ak[:, None] + bk[:, None] * ps[None, :]


def layer_logs(p):
    """ln(p_K+1/p_K) and alfa_K; the top layer (p_top = 0) has no log and alfa = ln 2"""
    lr = np.zeros(p[1:].shape, p.dtype)
    lr[1:] = np.log(p[2:] / p[1:-1])
    dp = p[1:] - p[:-1]
    alfa = np.empty(dp.shape, p.dtype)
    alfa[0] = F32LN2
    alfa[1:] = 1.0 - p[1:-1] / dp[1:] * lr[1:]
    return dp, lr, alfa


def geopotential(ak, bk, ps, t_td, phis):
    """full-level hydrostatic geopotential, top -> bottom"""
    p = p_int(ak, bk, ps)
    dp, lr, alfa = layer_logs(p)
    below = np.cumsum((RD * t_td * lr)[::-1], axis=0)[::-1]  # sum over J >= K
    below = np.concatenate([below[1:], np.zeros_like(below[:1])])  # sum over J > K
    return phis + below + alfa * RD * t_td


def cstep(f, ps, dps, *args_and_dirs):
    """directional derivative of f(ps, *args) along (dps, *dirs), by complex step"""
    args = [a + 1j * H * d for a, d in args_and_dirs]
    return np.imag(f(ps + 1j * H * dps, *args)) / H


def grad_lnp(ak, bk, ps, dps):
    """S&B (grad ln p)_K . e = (ln(p_K+1/p_K) grad p_K + alfa_K grad dp_K) / dp_K along
    the direction with d ps = dps (top -> bottom)"""
    p = p_int(ak, bk, ps)
    dp, lr, alfa = layer_logs(p)
    gp = cstep(lambda s: p_int(ak, bk, s), ps, dps)
    gdp = gp[1:] - gp[:-1]
    return (lr * gp[:-1] + alfa * gdp) / dp, gdp


def test_omega_continuity_and_energy_conversion():
    for kind, nl in (("sigma", 20), ("hybrid", 64), ("hybrid", 24)):
        ak, bk, lnps, tv, ug, vg, divg, d = columns(kind, nl, seed=nl)
        ps = np.exp(lnps)
        dlnpsdt, etadot, omg = O.omega_cols(ak, bk, lnps, ug, vg, divg, d["dlnpsdx"], d["dlnpsdy"])
        u, v, dv, t = ug[::-1], vg[::-1], divg[::-1], tv[::-1]      # top -> bottom
        p = p_int(ak, bk, ps)
        dp, lr, alfa = layer_logs(p)
        glx, gdpx = grad_lnp(ak, bk, ps, ps * d["dlnpsdx"])
        gly, gdpy = grad_lnp(ak, bk, ps, ps * d["dlnpsdy"])
        D = dv * dp + u * gdpx + v * gdpy                       # div(V_K dp_K)
        # column continuity
        np.testing.assert_allclose(ps * dlnpsdt, -D.sum(0), rtol=1e-12, atol=1e-12 * np.abs(D).sum(0).max())
        # each layer's mass budget, with etadot zero at the top and the surface
        assert np.all(etadot[0] == 0) and np.all(etadot[-1] == 0)
        ddpdt = np.diff(bk)[:, None] * ps * dlnpsdt
        resid = ddpdt + D + etadot[1:] - etadot[:-1]
        assert np.abs(resid).max() < 1e-12 * np.abs(D).max(), kind
        # omega/p per layer: V.(grad ln p) minus the layer's share of the mass
        # convergence above and within it
        above = np.concatenate([np.zeros_like(D[:1]), np.cumsum(D, 0)[:-1]])
        vgl = u * glx + v * gly
        want = vgl - (lr * above + alfa * D) / dp
        # the reference's top layer: advective part V.grad(dp_1)/dp_1 (no alfa_1),
        # and alfa_1 multiplies div_1 dp_1 + ps V_1.grad ln ps dB_1^2 (cb(1)*dbk(1)
        # with cb(1) = cg(1)*dbk(1), dyn_run.f90:349-350, where lower layers have cg*dbk)
        vgp1 = u[0] * gdpx[0] + v[0] * gdpy[0]                   # V_1.grad dp_1 = ps cg dB_1
        top = (vgp1 - alfa[0] * (dv[0] * dp[0] + vgp1 * (bk[1] - bk[0]))) / dp[0]
        sb_top = want[0].copy()
        want[0] = top
        got = omg[::-1]
        assert np.abs(got - want).max() < 1e-11 * np.abs(want).max(), kind
        # S&B energy conversion: sum_K R T dp (V.grad ln p - omega/p) = sum_K (phi - phi_s) D,
        # less the top layer's departure from the S&B form
        lhs = (RD * t * dp * (vgl - got)).sum(0)
        phi = geopotential(ak, bk, ps, t, 0.0)
        rhs = (phi * D).sum(0) - RD * t[0] * dp[0] * (top - sb_top)
        np.testing.assert_allclose(lhs, rhs, rtol=1e-9, atol=1e-10 * np.abs(rhs).max())


@pytest.mark.parametrize("kind,nl", [("sigma", 20), ("hybrid", 64), ("hybrid", 24)])
def test_presgrad_is_hydrostatic_gradient(kind, nl):
    ak, bk, lnps, tv, ug, vg, divg, d = columns(kind, nl, seed=100 + nl)
    ps = np.exp(lnps)
    px, py = O.presgrad_cols(ak, bk, lnps, tv, d["dtvdx"], d["dtvdy"], d["dphisdx"], d["dphisdy"],
                             d["dlnpsdx"], d["dlnpsdy"])
    t = tv[::-1]
    for got, dl, dT, dphis in ((px, d["dlnpsdx"], d["dtvdx"], d["dphisdx"]),
                               (py, d["dlnpsdy"], d["dtvdy"], d["dphisdy"])):
        gphi = cstep(lambda s, tt, ph: geopotential(ak, bk, s, tt, ph), ps, ps * dl,
                     (t, dT[::-1]), (d["phis"], dphis))
        gl, _ = grad_lnp(ak, bk, ps, ps * dl)
        want = -(gphi + RD * t * gl)
        assert np.abs(got[::-1] - want).max() < 1e-10 * np.abs(want).max(), kind
        if kind == "sigma":
            # p = sigma ps: (grad ln p)_K = grad ln ps, the textbook sigma-coordinate force
            np.testing.assert_allclose(gl[1:], np.broadcast_to(dl, gl[1:].shape), rtol=1e-12)
        # angular momentum: sum_K dp_K PGF_K = -grad(sum_K dp_K phi_K) + phi_s grad ps
        colphi = cstep(lambda s, tt, ph: (np.diff(p_int(ak, bk, s), axis=0) *
                                          geopotential(ak, bk, s, tt, ph)).sum(0),
                       ps, ps * dl, (t, dT[::-1]), (d["phis"], dphis))
        dp = np.diff(p_int(ak, bk, ps), axis=0)
        lhs = (dp * got[::-1]).sum(0)
        rhs = -colphi + d["phis"] * ps * dl
        np.testing.assert_allclose(lhs, rhs, rtol=1e-10, atol=1e-10 * np.abs(rhs).max())


@pytest.mark.parametrize("kind,nl", [("sigma", 20), ("hybrid", 64)])
def test_presgrad_isothermal_balance(kind, nl):
    """Over an isothermal column the continuous force is -grad phi_s - R T0 grad ln ps
    whatever the levels; S&B's discretisation keeps this exactly on every layer
    below the top, which carries alfa_1 = ln 2 in place of the exact log."""
    ak, bk, lnps, tv, ug, vg, divg, d = columns(kind, nl, seed=7)
    ps = np.exp(lnps)
    t0 = 255.0
    tv = np.full_like(tv, t0)
    z = np.zeros_like(tv)
    px, py = O.presgrad_cols(ak, bk, lnps, tv, z, z, d["dphisdx"], d["dphisdy"], d["dlnpsdx"], d["dlnpsdy"])
    p1 = ak[1] + bk[1] * ps
    for got, dl, dphis in ((px, d["dlnpsdx"], d["dphisdx"]), (py, d["dlnpsdy"], d["dphisdy"])):
        want = -dphis - RD * t0 * dl
        g = got[::-1]
        assert np.abs(g[1:] - want).max() < 1e-10 * np.abs(want).max()
        top = -dphis - RD * t0 * (dl + (F32LN2 - 1.0) * bk[1] * ps * dl / p1)
        assert np.abs(g[0] - top).max() < 1e-10 * np.abs(top).max()
