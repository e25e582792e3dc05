"""Stochastic physics on the GPU against the FP64 oracle, which draws from the
same counter RNG (member i of a GPU ensemble uses seed + i):

1. the spun-up svc, sppt, shum and addnoise patterns and their advance inside
   advance() (eager and HIP-graph steps continue the same streams), and the
   grids (sppt after its logit transform);
2. vorticity confinement modulated by svc in the dynamics tendencies;
3. additive streamfunction noise (1-2-1 filter, dissipation modulation) over
   three Held-Suarez steps, the forcing and the state;
4. a checkpoint taken with every scheme on continues bit-identically.
FP64 to 1e-9 (max-norm relative), FP32 within its tendency bound 2e-4.
"""
import numpy as np
import pytest

import cases
import levels
import oracle as O

pytestmark = pytest.mark.gpu

NLON, NLAT, T, NL, DT = 128, 64, 42, 20, 1200.0


def kind(name, amp, seed, tau=3 * 3600.0, lscale=5.0e5):
    return {name: amp, name + "_tau": tau, name + "_lscale": lscale, "iseed_" + name: seed}


ALL = dict(vcamp=6.0, **kind("svc", 0.6, 3), **kind("sppt", 0.5, 4), **kind("shum", 0.02, 5),
           **kind("addnoise", 2.0e5, 6), addnoise_vfilt=1, addnoise_dissfact=1)


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def pair(prec, nmem=1, member=0, **stoch):
    import gfsdc
    ak, bk = levels.hybrid_levels(NL)
    o = O.Model(NLON, NLAT, T, NL, ak, bk, DT)
    g = gfsdc.Context(NLON, NLAT, T, prec, nlev=NL, ak=ak, bk=bk, dt=DT, nmem=nmem)
    seeds = {k: v + member for k, v in stoch.items() if k.startswith("iseed")}
    o.stoch_init(**{**stoch, **seeds})
    g.stoch_init(**stoch)
    return o, g


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_patterns_match_oracle(prec):
    import gfsdc
    p = gfsdc.FP64 if prec == "fp64" else gfsdc.FP32
    print(f"stoch patterns {prec}", flush=True)
    o, g = pair(p, **ALL)
    tol = 1e-12 if prec == "fp64" else 1e-6
    for w in ("svc_spec", "sppt_spec", "shum_spec"):
        assert rel(g.stoch(w)[0], o.stoch(w)) < tol, w
    assert rel(g.stoch("addnoise")[0], o.stoch("addnoise")) < tol
    vrt, div, tv, lnps = cases.rest_state(o.sht.degree, o.sht.order, NL, seed=1, vort_rms=2e-5)
    o.set_state(vrt[0], div[0], tv[0], lnps[0])
    g.set_state(vrt, div, tv, lnps)
    o.advance(2)
    g.advance(2, use_graph=True)
    gtol = 1e-11 if prec == "fp64" else 1e-5
    for w in ("svc", "sppt", "shum"):
        assert rel(g.stoch(w)[0], o.stoch(w)) < gtol, w
    for w in ("svc_spec", "addnoise"):
        assert rel(g.stoch(w)[0], o.stoch(w)) < tol, w


def test_ensemble_member_seeds():
    import gfsdc
    print("stoch ensemble seeds", flush=True)
    st = dict(**kind("sppt", 0.5, 21), **kind("addnoise", 1e5, 30))
    o1, g = pair(gfsdc.FP64, nmem=2, member=1, **st)
    assert rel(g.stoch("sppt_spec")[1], o1.stoch("sppt_spec")) < 1e-12
    assert rel(g.stoch("addnoise")[1], o1.stoch("addnoise")) < 1e-12
    assert rel(g.stoch("sppt_spec")[0], o1.stoch("sppt_spec")) > 0.5


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_vorticity_confinement_tendencies(prec):
    import gfsdc
    p = gfsdc.FP64 if prec == "fp64" else gfsdc.FP32
    print(f"stoch vc {prec}", flush=True)
    o, g = pair(p, vcamp=8.0, **kind("svc", 0.7, 9))
    vrt, div, tv, lnps = cases.perturbed_state(o.sht.degree, o.sht.order, NL, seed=5)
    o.set_state(vrt[0], div[0], tv[0], lnps[0])
    g.set_state(vrt, div, tv, lnps)
    to, tg = o.getdyntend(1), g.getdyntend(1)
    tol = 1e-9 if prec == "fp64" else 2e-4
    for k in ("dvrt", "ddiv", "dtv", "dlnps"):
        assert rel(tg[k][0], to[k]) < tol, (k, rel(tg[k][0], to[k]))


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_additive_noise_steps(prec):
    import gfsdc
    p = gfsdc.FP64 if prec == "fp64" else gfsdc.FP32
    print(f"stoch addnoise {prec}", flush=True)
    o, g = pair(p, **ALL)
    vrt, div, tv, lnps = cases.rest_state(o.sht.degree, o.sht.order, NL, seed=2, vort_rms=2e-5)
    o.set_state(vrt[0], div[0], tv[0], lnps[0])
    g.set_state(vrt, div, tv, lnps)
    o.advance(3)
    g.advance(3, use_graph=True)
    so, sg = o.get_state(), g.get_state()
    if prec == "fp64":
        assert rel(g.stoch("psiforcing")[0], o.stoch("psiforcing")) < 1e-9
        for k in ("vrt", "div", "tv", "lnps"):
            assert rel(sg[k][0], so[k]) < 1e-9, (k, rel(sg[k][0], so[k]))
    else:
        for k in ("vrt", "div", "tv", "lnps"):
            e = np.linalg.norm(sg[k][0] - so[k]) / np.linalg.norm(so[k])
            assert e < 1e-5, (k, e)


def test_graph_replay_and_checkpoint_with_stochastic_physics():
    import gfsdc
    print("stoch graph + checkpoint fp32", flush=True)
    ak, bk = levels.hybrid_levels(NL)
    mk = lambda: gfsdc.Context(NLON, NLAT, T, gfsdc.FP32, nlev=NL, ak=ak, bk=bk, dt=DT, nmem=2)
    a, b, c = mk(), mk(), mk()
    for x in (a, b, c):
        x.stoch_init(**ALL)
        x.init_rest_state(300.0, 1.0e5, 2e-5, 1, 21, seed=8)
    a.advance(2, use_graph=True)
    b.advance(2, use_graph=False)
    for k in ("vrt", "div", "tv", "lnps"):
        assert np.array_equal(a.get_state()[k], b.get_state()[k]), k
    img = a.checkpoint()
    a.advance(3, use_graph=True)
    c.restore(img)
    c.advance(3, use_graph=True)
    for k in ("vrt", "div", "tv", "lnps"):
        assert np.array_equal(a.get_state()[k], c.get_state()[k]), k
    for w in ("svc_spec", "sppt_spec", "shum_spec", "addnoise", "psiforcing"):
        assert np.array_equal(a.stoch(w), c.stoch(w)), w
    # an image with other schemes on is refused
    d = mk()
    with pytest.raises(gfsdc.GfsdcError, match="checkpoint"):
        d.restore(img)
