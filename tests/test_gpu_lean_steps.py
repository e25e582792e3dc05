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
    for name in ("ug", "vg", "vrtg", "divg", "virtempg", "dlnpdtg", "etadot", "psg", "dlnpsdt"):
        a, b = g.grid(name)[0], o.grid(name)
        err = np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)
        assert err < 1e-9, (name, err)
