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

NLON, NLAT, T = 64, 32, 21
SQ = np.sqrt(4 * np.pi)


def f32(x):
    return float(np.float32(x))


# IMEX tableau of semimp_data.f90:74-88 (default-real literals); its order and
# consistency identities are asserted in test_oracle_dynamics.py
A = {2: (f32(1 / 3),), 3: (f32(1 / 6), f32(1 / 2)), 4: (f32(1 / 2), f32(-1 / 2), 1.0)}
AA = {2: (f32(1 / 6),), 3: (f32(1 / 3), 0.0), 4: (f32(3 / 8), 0.0, f32(3 / 8))}
AI = {2: f32(1 / 6), 3: f32(1 / 3), 4: f32(1 / 4)}


def deep_levels(nl, p0=1.0e5):
    """hybrid levels crowded toward the top (eta^3 spacing), so that the top
    layer lies above sigma = 0.002 where the Rayleigh sponge of setdampspec acts"""
    eta = np.linspace(0.0, 1.0, nl + 1) ** 3
    bk = np.clip((eta - 0.2) / 0.8, 0.0, None) ** 1.5
    bk[-1] = 1.0
    ak = p0 * (eta - bk)
    ak[-1] = 0.0
    return ak, bk


def make(nl, dt, deep=False, **kw):
    ak, bk = deep_levels(nl) if deep else levels.hybrid_levels(nl)
    return O.Model(NLON, NLAT, T, nl, ak, bk, dt, physics=0, **kw)


def lin(sm, lap, y):
    """L(y) on (vrt, div, tv, lnps): the semi-implicit linear terms (run.f90:326-333)"""
    vrt, div, tv, lnps = y
    return (np.zeros_like(vrt),
            -lap[None, :] * (sm["amhyb"] @ tv + sm["tor_hyb"][:, None] * lnps[None, :]),
            -(sm["bmhyb"] @ div),
            -(sm["svhyb"] @ div))


def tend(m, y, stage):
    m.set_state(*y)
    o = m.getdyntend(stage)
    return (o["dvrt"], o["ddiv"], o["dtv"], o["dlnps"])


def implicit_solve(sm, lap, c, dt, rhs):
    """solve (I - c dt L) x = rhs for (div, tv, lnps), one dense system per coefficient"""
    nl = sm["amhyb"].shape[0]
    nlm = lap.size
    M = np.zeros((nlm, 2 * nl + 1, 2 * nl + 1))
    h = c * dt
    M[:, :nl, nl:2 * nl] = h * lap[:, None, None] * sm["amhyb"][None]
    M[:, :nl, 2 * nl] = h * lap[:, None] * sm["tor_hyb"][None, :]
    M[:, nl:2 * nl, :nl] = h * sm["bmhyb"][None]
    M[:, 2 * nl, :nl] = h * sm["svhyb"][None, :]
    M += np.eye(2 * nl + 1)[None]
    _, div, tv, lnps = rhs
    b = np.concatenate([div, tv, lnps[None, :]], 0).T[:, :, None]      # (nlm, 2nl+1, 1)
    x = np.linalg.solve(M + 0j, b)[:, :, 0].T
    return x[:nl], x[nl:2 * nl], x[2 * nl]


def imex_step(m, y0, damp=True):
    sm, dt, lap = m.semimp(), m.cfg.dt, m.sht.lap
    Ys, Ns, Ls = [y0], [], []
    for s in (2, 3, 4):
        Y = Ys[-1]
        F = tend(m, Y, s - 1)
        Ly = lin(sm, lap, Y)
        Ns.append(tuple(f - l for f, l in zip(F, Ly)))
        Ls.append(Ly)
        rhs = []
        for q in range(4):
            acc = y0[q].copy()
            for r in range(s - 1):
                acc = acc + dt * (A[s][r] * Ns[r][q] + AA[s][r] * Ls[r][q])
            rhs.append(acc)
        div, tv, lnps = implicit_solve(sm, lap, AI[s], dt, rhs)
        Ys.append((rhs[0], div, tv, lnps))
    vrt, div, tv, lnps = Ys[-1]
    if damp:
        # run.f90:493-502: implicit (backward Euler) damping, du/dt = -(nu + r) u,
        # with the hyperdiffusion rate nu = -disspec diff_prof >= 0 on every field
        # and the Rayleigh sponge rate r = dmp_prof >= 0 on vorticity and divergence
        d = m.damp()
        nu = -d["disspec"][None, :] * d["diff_prof"][:, None]
        r = d["dmp_prof"][:, None]
        assert nu.min() >= 0 and r.min() >= 0
        vrt, div = vrt / (1.0 + (nu + r) * dt), div / (1.0 + (nu + r) * dt)
        tv = tv / (1.0 + nu * dt)
    return vrt, div, tv, lnps


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
