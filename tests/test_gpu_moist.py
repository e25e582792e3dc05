"""GPU parity of the moist parts of the dycore against the FP64 oracle:
tracer advection inside getdyntend (dyn_run.f90:279-291 with the Thuburn
limiter of getvadv_tracers), the dry-mass diagnostic with tracer mass
(dyn_run.f90:135-153), the simple moist physics (phy_run_simple.f90 around
simple_physics_v5_1.f90) and full advance steps with physics and the mass
fixer using pwat.

Tolerances: FP64 path 1e-9 relative max-norm. FP32 path: tendencies 2e-4
relative max-norm; simple physics 2e-3 relative L2 (the saturation test
q > qsat flips at a few points when inputs differ in the last FP32 bit, which
moves single grid points, not the field).
"""
import numpy as np
import pytest

import cases
import levels
import oracle as O

pytestmark = pytest.mark.gpu

NLON, NLAT, T = 128, 64, 42


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def rel2(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def moist_pair(precision, nlev=20, ntrac=1, dt=1200.0, **kw):
    import gfsdc
    ak, bk = levels.hybrid_levels(nlev)
    o = O.Model(NLON, NLAT, T, nlev, ak, bk, dt, ntrac=ntrac, **kw)
    g = gfsdc.Context(NLON, NLAT, T, precision, nlev=nlev, ak=ak, bk=bk, dt=dt, ntrac=ntrac, **kw)
    vrt, div, tv, lnps = cases.perturbed_state(o.sht.degree, o.sht.order, nlev, seed=11, amp=0.5)
    q = o.sht.grdtospec(cases.moist_grid(o.sht.lats, NLON, ak, bk, seed=3))
    trac = np.stack([q] + [0.5 * q + (1e-6 * np.sqrt(4 * np.pi)) * (o.sht.degree == 0)
                           for _ in range(ntrac - 1)])
    o.set_state(vrt[0], div[0], tv[0], lnps[0], trac=trac)
    g.set_state(vrt, div, tv, lnps, trac=trac[None])
    return o, g


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("ntrac", [1, 2])
def test_getdyntend_tracers(prec, ntrac):
    import gfsdc
    p = gfsdc.FP64 if prec == "fp64" else gfsdc.FP32
    o, g = moist_pair(p, ntrac=ntrac, physics=0)
    to, tg = o.getdyntend(1), g.getdyntend(1)
    tol = 1e-9 if prec == "fp64" else 2e-4
    for k in ("dvrt", "ddiv", "dtv", "dlnps"):
        assert rel(tg[k][0], to[k]) < tol, (k, rel(tg[k][0], to[k]))
    for nt in range(ntrac):
        e = rel(tg["dtrac"][0, nt], to["dtrac"][nt])
        assert e < tol, (nt, e)
    # pdryini = mean ps - sum over tracers of the mean column tracer mass
    assert abs(g.pdryini[0] - o.pdryini) < (1e-9 if prec == "fp64" else 2e-6) * o.pdryini
    assert rel(g.grid("tracerg")[0], o.grid("tracerg")) < (1e-9 if prec == "fp64" else 1e-5)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("sst,turb", [(0, 1), (1, 1), (2, 1), (0, 0)])
def test_simple_phytend(prec, sst, turb):
    import gfsdc
    p = gfsdc.FP64 if prec == "fp64" else gfsdc.FP32
    o, g = moist_pair(p, physics=2, simple_test=sst, turbflux=turb)
    o.getdyntend(3, inverse_only=True)
    g.getdyntend(3, inverse_only=True)
    to, tg = o.simple_phytend(), g.simple_phytend()
    for k in ("dvrt", "ddiv", "dtv"):
        e = rel(tg[k][0], to[k]) if prec == "fp64" else rel2(tg[k][0], to[k])
        assert e < (1e-9 if prec == "fp64" else 2e-3), (k, e)
    e = rel(tg["dtrac"][0, 0], to["dtrac"][0]) if prec == "fp64" else rel2(tg["dtrac"][0, 0], to["dtrac"][0])
    assert e < (1e-9 if prec == "fp64" else 2e-3), ("dtrac", e)
    for name in ("precip", "pwat"):
        a, b = g.grid(name)[0, 0], o.grid(name)[0]
        e = rel(a, b) if prec == "fp64" else rel2(a, b)
        assert e < (1e-9 if prec == "fp64" else 2e-3), (name, e)
    assert o.grid("precip").max() > 0  # the case condenses


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_advance_simple_physics(prec):
    """two full steps: IMEX RK3 with tracers, damping, simple physics and the
    mass fixer with pwat (run.f90:509-538, dyn_run.f90:605-634)"""
    import gfsdc
    p = gfsdc.FP64 if prec == "fp64" else gfsdc.FP32
    o, g = moist_pair(p, physics=2, simple_test=2, turbflux=1, massfix=1)
    o.advance(2)
    g.advance(2, use_graph=True)
    so, sg = o.get_state(), g.get_state()
    tol = 1e-9 if prec == "fp64" else 1e-4
    for k in ("vrt", "div", "tv", "lnps"):
        e = rel2(sg[k][0], so[k])
        assert e < tol, (k, e)
    e = rel2(sg["trac"][0], so["trac"])
    assert e < tol, ("trac", e)
    assert abs(g.pdryini[0] - o.pdryini) < 1e-9 * o.pdryini if prec == "fp64" else True


def test_hs_with_passive_tracer_keeps_tracer_physics_free():
    """Held-Suarez with a tracer: the physics leaves the tracer to the dynamics."""
    import gfsdc
    o, g = moist_pair(gfsdc.FP64, physics=1, massfix=1)
    o.advance(1)
    g.advance(1)
    so, sg = o.get_state(), g.get_state()
    for k in ("vrt", "div", "tv", "lnps"):
        assert rel2(sg[k][0], so[k]) < 1e-9, k
    assert rel2(sg["trac"][0], so["trac"]) < 1e-9
