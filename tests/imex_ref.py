"""Reference IMEX-RK step for the oracle tests: the generic additive
Runge-Kutta method with the tableau of semimp_data.f90:74-88, its implicit
part solved as one dense linear system per spectral coefficient."""
import numpy as np

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


def imex_step(m, y0, damp=True, force=None):
    """one step of the generic IMEX-RK method; force (vrt, div, tv, lnps), if
    given, is a tendency held constant over the step and added to every stage"""
    sm, dt, lap = m.semimp(), m.cfg.dt, m.sht.lap
    Ys, Ns, Ls = [y0], [], []
    for s in (2, 3, 4):
        Y = Ys[-1]
        F = tend(m, Y, s - 1)
        if force is not None:
            F = tuple(f + g for f, g in zip(F, force))
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
