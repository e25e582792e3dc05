"""The lean Held-Suarez physics pass of the steps before an advance's last.

advance(n) with Held-Suarez physics synthesises, in the physics pass of its
first n - 1 steps, only the grids the forcing reads (vorticity, divergence,
temperature, ln ps: phy_run_hs.f90:66-76) and forms ps without the getomega
column pass. The state must be bit-identical to n single steps, each of which
runs the full pass, and the last step must leave the same grid fields and
run.f90:157-162 diagnostics. Checked in FP64 and FP32, eager and HIP-graph,
with two members, the dry-mass fixer and the stochastic temperature pattern
(whose forcing acts on the spectral state between the passes).
"""
import numpy as np
import pytest

import cases
import levels

pytestmark = pytest.mark.gpu

NLON, NLAT, T, NL = 128, 64, 42, 20
GRIDS = ("ug", "vg", "vrtg", "divg", "virtempg", "dlnpdtg", "etadot", "psg", "dlnpsdt", "dlnpsdx",
         "dlnpsdy")


def ctx(prec, **kw):
    import gfsdc
    ak, bk = levels.hybrid_levels(NL)
    c = gfsdc.Context(NLON, NLAT, T, prec, nlev=NL, ak=ak, bk=bk, dt=1200.0, nmem=2,
                      tpat_amp=0.5, tpat_tau=6 * 3600.0, tpat_lscale=5.0e5, tpat_seed=7, **kw)
    c.init_rest_state(300.0, 1.0e5, 2e-5, 1, 21, seed=3)
    return c


@pytest.mark.parametrize("graph", [False, True])
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_lean_steps_are_bit_identical(prec, graph):
    import gfsdc
    p = gfsdc.FP64 if prec == "fp64" else gfsdc.FP32
    print(f"lean steps {prec} graph={graph}", flush=True)
    a, b = ctx(p), ctx(p)
    a.advance(5, use_graph=graph)          # 4 lean steps, then a full one
    for _ in range(5):
        b.advance(1, use_graph=graph)      # full steps only
    # the lean passes formed the drag on the lowest levels only (6 or 7 of these
    # 20 near ps = 1e5), the full passes on all of them
    assert 0 < a.hs_drag_levels() < NL
    assert b.hs_drag_levels() == 0
    sa, sb = a.get_state(), b.get_state()
    for k in ("vrt", "div", "tv", "lnps"):
        assert np.array_equal(sa[k], sb[k]), k
    assert np.array_equal(a.pattern(), b.pattern())
    for g in GRIDS:
        assert np.array_equal(a.grid(g), b.grid(g)), g
    da, db = a.step_diag(), b.step_diag()
    for k in da:
        assert np.array_equal(da[k], db[k]), k
    assert a.time == b.time == 5 * 1200.0


def phillips_sigma(ak, bk, ps, kappa=2.0 / 7.0):
    """p_k / ps of every level (level 0 the lowest) from the Phillips layer
    pressure of pressure_data.f90:87-88, in FP64"""
    pi = ak + bk * ps
    p1, p2 = pi[:-1], pi[1:]
    with np.errstate(divide="ignore", invalid="ignore"):
        p = ((p2 ** (kappa + 1) - p1 ** (kappa + 1)) / ((kappa + 1) * (p2 - p1))) ** (1 / kappa)
    return p[::-1] / ps


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_drag_levels_cover_the_boundary_layer(prec):
    """On sigma levels p / ps does not depend on ps, so the levels with drag
    (p / ps > sigbot = 0.7, phy_run_hs.f90:52-58) are fixed: the lowest 6 of
    the 20 (sigma 0.975 .. 0.725). The lean pass may take a level more (its
    test uses the layer's lower interface), never fewer; the state still
    matches full steps bit for bit."""
    import gfsdc
    p = gfsdc.FP64 if prec == "fp64" else gfsdc.FP32
    print(f"drag levels {prec}", flush=True)
    ak, bk = levels.levels_for("T42L20")
    assert np.all(ak == 0)
    sig = phillips_sigma(ak, bk, 1.0e5)
    nbl = int(np.sum(sig > 0.7))
    assert nbl == 6 and np.all(sig[:nbl] > 0.7)

    def mk():
        c = gfsdc.Context(NLON, NLAT, T, p, nlev=NL, ak=ak, bk=bk, dt=1200.0)
        c.init_rest_state(300.0, 1.0e5, 2e-5, 1, 21, seed=5)
        return c
    a, b = mk(), mk()
    a.advance(3, use_graph=True)
    for _ in range(3):
        b.advance(1, use_graph=True)
    assert nbl <= a.hs_drag_levels() <= nbl + 1
    sa, sb = a.get_state(), b.get_state()
    for k in ("vrt", "div", "tv", "lnps"):
        assert np.array_equal(sa[k], sb[k]), k


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_drag_reaching_more_levels_than_the_last_pass(prec):
    """The lean pass first synthesises the vorticity and divergence of the
    levels where the previous lean pass found drag, plus one; where the drag
    now reaches higher it synthesises the rest before the forcing. Halving ps
    between two advances (7 levels with drag near 1000 hPa on these hybrid
    levels, 12 near 500 hPa) takes that path; the state still matches full
    steps bit for bit, through the mass fixer's ln ps update as well."""
    import gfsdc
    p = gfsdc.FP64 if prec == "fp64" else gfsdc.FP32
    print(f"drag levels rising {prec}", flush=True)
    ak, bk = levels.hybrid_levels(NL)

    def mk():
        c = gfsdc.Context(NLON, NLAT, T, p, nlev=NL, ak=ak, bk=bk, dt=1200.0)
        c.init_rest_state(300.0, 1.0e5, 2e-5, 1, 21, seed=6)
        return c
    a, b = mk(), mk()
    a.advance(3, use_graph=True)
    for _ in range(3):
        b.advance(1, use_graph=True)
    n0 = a.hs_drag_levels()
    assert n0 in (6, 7, 8)
    for c in (a, b):
        st = c.get_state()
        st["lnps"][:, 0] -= np.log(2.0) * np.sqrt(4 * np.pi)  # the (0, 0) coefficient: ps / 2
        c.set_state(st["vrt"], st["div"], st["tv"], st["lnps"])
    # one lean step (the dry-mass fixer then restores ps) and a full one
    a.advance(2, use_graph=True)
    for _ in range(2):
        b.advance(1, use_graph=True)
    assert a.hs_drag_levels() >= n0 + 2
    sa, sb = a.get_state(), b.get_state()
    for k in ("vrt", "div", "tv", "lnps"):
        assert np.array_equal(sa[k], sb[k]), k


def test_profiled_lean_step_refuses_the_grids():
    import gfsdc
    print("lean profile_step", flush=True)
    a, b = ctx(gfsdc.FP32), ctx(gfsdc.FP32)
    names = [n for n, _ in a.profile_step()]
    assert "psg" in names and "columns_inverse" not in names
    b.advance(2, use_graph=False)
    a.advance(1, use_graph=False)
    assert np.array_equal(a.get_state()["vrt"], b.get_state()["vrt"])
    # the grids are complete again after the full step
    assert np.array_equal(a.grid("ug"), b.grid("ug"))
    a.profile_step()
    with pytest.raises(gfsdc.GfsdcError, match="lean"):
        a.grid("ug")
    with pytest.raises(gfsdc.GfsdcError, match="lean"):
        a.step_diag()
    a.getdyntend(1, inverse_only=True)
    a.grid("ug")


@pytest.mark.parametrize("physics", [0, 1])
def test_grids_after_advance_match_oracle(physics):
    """the grid fields an advance leaves (grid_data / pressure_data after the
    last getdyntend of its last step: the physics pass's with Held-Suarez, the
    third dynamics pass's without physics, where the column kernel stores the
    omega/p and etadot grids itself) against the FP64 oracle after the same
    steps, to 1e-9"""
    import gfsdc
    import oracle as O
    print(f"grids after advance physics={physics}", flush=True)
    ak, bk = levels.hybrid_levels(NL)
    o = O.Model(NLON, NLAT, 42, NL, ak, bk, 1200.0, physics=physics)
    g = gfsdc.Context(NLON, NLAT, 42, gfsdc.FP64, nlev=NL, ak=ak, bk=bk, dt=1200.0, physics=physics)
    vrt, div, tv, lnps = cases.rest_state(o.sht.degree, o.sht.order, NL, seed=4, vort_rms=2e-5)
    o.set_state(vrt[0], div[0], tv[0], lnps[0])
    g.set_state(vrt, div, tv, lnps)
    o.advance(3)
    g.advance(3, use_graph=True)
    # dlnpsdx is formed in the Fourier synthesis from ln ps's Legendre sums (i m / a),
    # dlnpsdy by a Legendre synthesis of its own
    for name in ("ug", "vg", "vrtg", "divg", "virtempg", "dlnpdtg", "etadot", "psg", "dlnpsdt", "dlnpsdx",
                 "dlnpsdy"):
        a, b = g.grid(name)[0], o.grid(name)
        err = np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)
        assert err < 1e-9, (name, err)
