"""GPU parity of the AR(1) temperature pattern (patterngenerator.f90 as
driven by stoch_data.f90 / run.f90) against the FP64 oracle, which draws
from the same counter RNG: the spun-up pattern, its per-step advance inside
advance() (HIP graph replays continue the random stream), the forcing it
applies, and the ensemble folding (member i = seed + i)."""
import numpy as np
import pytest

import cases
import levels
import oracle as O

pytestmark = pytest.mark.gpu

NLON, NLAT, T, NL, DT = 128, 64, 42, 20, 1200.0
PAT = dict(tpat_amp=0.5, tpat_tau=6 * 3600.0, tpat_lscale=5.0e5)


def rel2(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def oracle_model(seed, **kw):
    ak, bk = levels.hybrid_levels(NL)
    return O.Model(NLON, NLAT, T, NL, ak, bk, DT, tpat_seed=seed, **PAT, **kw)


def gpu_context(prec, seed, nmem=1, **kw):
    import gfsdc
    ak, bk = levels.hybrid_levels(NL)
    return gfsdc.Context(NLON, NLAT, T, prec, nlev=NL, ak=ak, bk=bk, dt=DT, nmem=nmem, tpat_seed=seed,
                         **PAT, **kw)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_spun_up_pattern_matches_oracle(prec):
    import gfsdc
    p = gfsdc.FP64 if prec == "fp64" else gfsdc.FP32
    o = oracle_model(5)
    g = gpu_context(p, 5)
    e = rel2(g.pattern()[0], o.pattern()["spec"])
    assert e < (1e-12 if prec == "fp64" else 2e-5), e


def test_ensemble_members_follow_their_seeds():
    import gfsdc
    g = gpu_context(gfsdc.FP64, 11, nmem=3)
    pg = g.pattern()
    for i in range(3):
        assert rel2(pg[i], oracle_model(11 + i).pattern()["spec"]) < 1e-12
    assert rel2(pg[0], pg[1]) > 0.5  # independent members


@pytest.mark.parametrize("use_graph", [False, True])
def test_advance_with_pattern_forcing(use_graph):
    """two Held-Suarez steps with the pattern advanced each step and applied as
    dt * vfact * pattern / tau to the temperature"""
    import gfsdc
    o = oracle_model(7, physics=1)
    g = gpu_context(gfsdc.FP64, 7, physics=1)
    vrt, div, tv, lnps = cases.perturbed_state(o.sht.degree, o.sht.order, NL, seed=4, amp=0.3)
    o.set_state(vrt[0], div[0], tv[0], lnps[0])
    g.set_state(vrt, div, tv, lnps)
    o.advance(2)
    g.advance(2, use_graph=use_graph)
    assert rel2(g.pattern()[0], o.pattern()["spec"]) < 1e-12
    so, sg = o.get_state(), g.get_state()
    for k in ("vrt", "div", "tv", "lnps"):
        assert rel2(sg[k][0], so[k]) < 1e-9, k
