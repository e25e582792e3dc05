"""Pins the oracle's advance (run.f90:194-565: the IMEX RK3 sequence with the
semi-implicit solve of semimp_data.f90) by two checks that do not reuse its
algebra.

1. One step equals the generic additive (IMEX) Runge-Kutta method applied
   stage by stage. The numpy step here takes the tendency F from the oracle's
   getdyntend and the linear operator L from the semi-implicit operators
   (pinned elsewhere by the linearisation test), sets N = F - L, and forms
   each stage as
       Y_s = y0 + dt sum_{r<s} (a_sr N(Y_r) + aa_sr L(Y_r)) + dt aa_ss L(Y_s),
   solving the implicit part as one dense linear system per coefficient over
   (divergence, temperature, ln ps). The reference instead removes the linear
   terms from the stage tendency, solves for divergence alone with the
   precomputed d_hyb_m(:,:,n+1,stage) and back-substitutes; those are
   equivalent only if the tableau rows, the stage inverse and the
   back-substitution are all right.
2. Order of accuracy: on small-amplitude gravity waves about a resting
   isothermal state without rotation the dynamics is linear, dy/dt = J y, with
   J probed from getdyntend by central differences (exact for the quadratic
   nonlinearity). The exact solution is expm(J t) y0 per spectral coefficient;
   halving dt must cut the error of advance by about 4 (second order), with
   the semi-implicit reference state (300 K, 800 hPa) differing from the base
   state (1000 hPa) so that an explicit linear part is present too.
"""
import numpy as np
import pytest
import scipy.linalg

import cases
import levels
import oracle as O
from imex_ref import NLON, NLAT, T, SQ, deep_levels, f32, imex_step, make, tend


def test_stratospheric_sponge_closed_form():
    """setdampspec (dyn_init.f90:480-491): r_k = ln(0.002/sigma_k) / taustratdamp
    on layers above sigma = 0.002 (sigma_k the layer mid-point of the interface
    sigmas A/101300 + B), zero below"""
    m = make(12, 1200.0, deep=True)
    ak, bk = deep_levels(12)
    si = (ak / np.float32(101300.0) + bk)[::-1]              # bottom -> top
    sl = 0.5 * (si[:-1] + si[1:])
    want = np.where(sl < f32(0.002), np.log(f32(0.002) / sl) / f32(5 * 86400.0), 0.0)
    got = m.damp()["dmp_prof"]
    assert (want > 0).sum() >= 1 and got[-1] > 0
    np.testing.assert_allclose(got, want, rtol=1e-6, atol=0)


def test_advance_is_the_imex_tableau_stage_by_stage():
    nl = 12
    m = make(nl, 1200.0, deep=True)
    s = m.sht
    vrt, div, tv, lnps = cases.perturbed_state(s.degree, s.order, nl, seed=11, amp=1.0)
    y0 = (vrt[0], div[0], tv[0], lnps[0])
    want = imex_step(m, y0)
    m.set_state(*y0)
    m.advance(1)
    got = m.get_state()
    for name, w in zip(("vrt", "div", "tv", "lnps"), want):
        # degree 0 is left out: for vrt and div it is rounding noise about zero
        # (lap = 0), for tv and ln ps a mean that the step hardly changes
        g, w = got[name][..., 1:], w[..., 1:]
        err = np.abs(g - w).max() / np.abs(w).max()
        assert err < 1e-11, (name, err)
    # and the step did something beyond its first stage: the explicit and
    # implicit parts both carry weight
    assert np.abs(got["div"] - y0[1]).max() > 1e-3 * np.abs(y0[1]).max()


def _gravity_wave_setup(nl):
    m = make(nl, 600.0, omega=0.0, efold=1e30)
    s = m.sht
    rng = np.random.default_rng(3)
    sel = (s.degree >= 1) & (s.degree <= 8)

    def rnd(scale):
        x = (rng.standard_normal(m.nlm) + 1j * rng.standard_normal(m.nlm)) * scale
        x[s.order == 0] = x[s.order == 0].real
        x[~sel] = 0
        return x

    base_tv = np.zeros((nl, m.nlm), np.complex128)
    base_tv[:, 0] = 300.0 * SQ
    base_lnps = np.zeros(m.nlm, np.complex128)
    base_lnps[0] = np.log(1.0e5) * SQ
    zero = np.zeros((nl, m.nlm), np.complex128)
    y0 = (zero, np.array([rnd(1e-9) for _ in range(nl)]), np.array([rnd(1e-3) for _ in range(nl)]),
          rnd(1e-6))
    # Jacobian columns by central differences, one vertical component at a time,
    # every coefficient at once (the linear problem does not couple coefficients)
    c = np.where(sel, rnd(1.0), 0)
    cs = np.where(sel, c, 1.0)
    ncomp = 2 * nl + 1
    J = np.zeros((m.nlm, ncomp, ncomp), np.complex128)
    for j in range(ncomp):
        pert = [zero.copy(), zero.copy(), np.zeros(m.nlm, np.complex128)]
        eps = (1e-9, 1e-3, 1e-6)[0 if j < nl else 1 if j < 2 * nl else 2]
        if j < nl:
            pert[0][j] = eps * c
        elif j < 2 * nl:
            pert[1][j - nl] = eps * c
        else:
            pert[2] = eps * c
        out = []
        for sgn in (1.0, -1.0):
            y = (zero, sgn * pert[0], base_tv + sgn * pert[1], base_lnps + sgn * pert[2])
            f = tend(m, y, 2)
            out.append(np.concatenate([f[1], f[2], f[3][None, :]], 0))
        J[:, :, j] = ((out[0] - out[1]) / (2 * eps) / cs[None, :]).T
    return m, sel, base_tv, base_lnps, y0, J


def test_imex_converges_at_second_order_on_gravity_waves():
    nl = 10
    m, sel, base_tv, base_lnps, y0, J = _gravity_wave_setup(nl)
    tend_total = 7200.0
    x0 = np.concatenate([y0[1], y0[2], y0[3][None, :]], 0).T          # (nlm, 2nl+1)
    exact = np.zeros_like(x0)
    for k in np.nonzero(sel)[0]:
        exact[k] = scipy.linalg.expm(J[k] * tend_total) @ x0[k]
    # per-component scaling, so div, T and ln ps count alike in the norms
    w = np.concatenate([np.full(nl, 1e-9), np.full(nl, 1e-3), [1e-6]])
    # the solution moved well away from the start (the fastest wave, n = 8, has a
    # period of about 3.9 h)
    assert np.linalg.norm((exact - x0)[sel] / w) > 0.5 * np.linalg.norm(x0[sel] / w)
    errs = []
    ak, bk = levels.hybrid_levels(nl)
    for dt in (1200.0, 600.0, 300.0):
        mm = O.Model(NLON, NLAT, T, nl, ak, bk, dt, physics=0, omega=0.0, efold=1e30)
        mm.set_state(y0[0], y0[1], base_tv + y0[2], base_lnps + y0[3])
        mm.advance(int(round(tend_total / dt)))
        st = mm.get_state()
        x = np.concatenate([st["div"], st["tv"] - base_tv, (st["lnps"] - base_lnps)[None, :]], 0).T
        errs.append(np.linalg.norm((x - exact)[sel] / w) / np.linalg.norm(exact[sel] / w))
        # no rotation: vorticity stays at rounding level (typical values are 1e-5)
        assert np.abs(st["vrt"]).max() < 1e-11
    r1, r2 = errs[0] / errs[1], errs[1] / errs[2]
    assert 3.0 < r1 < 5.0 and 3.3 < r2 < 4.7, errs
    assert errs[2] < 1e-2, errs
