"""Restart, digital filter initialisation and IAU on the GPU.

1. Device checkpoint: a run that saves its image after k steps and goes on is
   bit-identical to a fresh context that loads the image and runs the same
   steps (FP64 and FP32, with the stochastic pattern and two members, eager
   and HIP-graph steps); a context of another shape refuses the image.
2. Sigma-file restart (readin_sig / wrtout_sig, dyn_init.f90:120-191): the
   GPU writes a file, a new GPU context reads it and runs on; the oracle runs
   from the same file (its transforms do its ln ps conversion) and the two
   agree to the FP64 bar; the restart also stays within FP32 file rounding of
   the run that did not stop.
3. DFI (run.f90:58-145) and IAU (iau_module.f90, run.f90:250-436) against the
   oracle on the same inputs: FP64 to 1e-9 (max-norm relative), FP32 to 1e-5
   (2-norm relative, the FP32 bound of the one-day advance test).
"""
import numpy as np
import pytest

import cases
import levels
import oracle as O
import sigio

pytestmark = pytest.mark.gpu

NLON, NLAT, T, NL = 128, 64, 42, 20
PAT = dict(tpat_amp=0.5, tpat_tau=6 * 3600.0, tpat_lscale=5.0e5, tpat_seed=3)


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def err(a, b, prec):
    """FP64: max-norm relative error; FP32: 2-norm relative error"""
    if prec == "fp64":
        return rel(a, b)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def ctx(prec, nmem=1, dt=1200.0, **kw):
    import gfsdc
    ak, bk = levels.hybrid_levels(NL)
    return gfsdc.Context(NLON, NLAT, T, prec, nlev=NL, ak=ak, bk=bk, dt=dt, nmem=nmem, **kw)


def oracle(dt=1200.0, **kw):
    ak, bk = levels.hybrid_levels(NL)
    return O.Model(NLON, NLAT, T, NL, ak, bk, dt, **kw)


def topo_spec(sht_degree, sht_order, seed=1):
    rng = np.random.default_rng(seed)
    n = np.maximum(sht_degree, 1)
    t = (rng.standard_normal(n.size) + 1j * rng.standard_normal(n.size)) * 800.0 / n ** 1.5
    t[sht_order == 0] = t[sht_order == 0].real
    t[(sht_degree > 30)] = 0
    return t


@pytest.mark.parametrize("graph", [False, True])
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_checkpoint_continuation_is_bit_identical(prec, graph):
    import gfsdc
    p = gfsdc.FP64 if prec == "fp64" else gfsdc.FP32
    print(f"checkpoint {prec} graph={graph}", flush=True)
    a = ctx(p, nmem=2, **PAT)
    a.init_rest_state(300.0, 1.0e5, 2e-6, 1, 21, seed=4)
    a.set_topo(topo_spec(a.degree, a.order))
    a.time = 7200.0
    a.advance(3, use_graph=graph)
    img = a.checkpoint()
    a.advance(4, use_graph=graph)
    sa, pa = a.get_state(), a.pattern()
    b = ctx(p, nmem=2, **PAT)
    b.restore(img)
    assert b.time == 7200.0 + 3 * 1200.0
    assert np.array_equal(b.pdryini, img_pdry(a, img))
    b.advance(4, use_graph=graph)
    sb, pb = b.get_state(), b.pattern()
    for k in ("vrt", "div", "tv", "lnps"):
        assert np.array_equal(sa[k], sb[k]), k
    assert np.array_equal(pa, pb)
    assert a.time == b.time == 7200.0 + 7 * 1200.0
    # the restored orography is the image's, at the context's precision
    c = np.complex128 if p == gfsdc.FP64 else np.complex64
    np.testing.assert_array_equal(b.topo(), a.topo().astype(c).astype(np.complex128))
    # another shape, precision or pattern setting refuses the image
    for other in (ctx(p, nmem=1, **PAT), ctx(p, nmem=2), ctx(gfsdc.FP32 if p == gfsdc.FP64 else gfsdc.FP64,
                                                                  nmem=2, **PAT)):
        with pytest.raises(gfsdc.GfsdcError, match="checkpoint"):
            other.restore(img)
    with pytest.raises(gfsdc.GfsdcError, match="magic"):
        b.restore(np.zeros_like(img))


def img_pdry(c, img):
    """pdryini section of a checkpoint image (after header, state, orography)"""
    import gfsdc
    isz = 16 if c.precision == gfsdc.FP64 else 8
    nfs = c.nlev * (3 + c.ntrac) + 1
    off = 128 + isz * c.nlm * (nfs + 1) * c.nmem
    return img[off:off + 8 * c.nmem].view(np.float64)


def test_sigma_file_restart(tmp_path):
    import gfsdc
    print("sigma restart fp64", flush=True)
    g = ctx(gfsdc.FP64)
    g.init_rest_state(300.0, 1.0e5, 2e-6, 1, 21, seed=6)
    g.set_topo(topo_spec(g.degree, g.order, seed=2))
    g.advance(3)
    path = tmp_path / "SIG.F01"
    g.write_sigma(path)
    head, data = sigio.read(path)
    assert head["fhour"] == 1.0 and head["jcap"] == T and head["levs"] == NL
    # the run that does not stop
    g.advance(3)
    cont = g.get_state()
    # GPU restart from the file
    r = ctx(gfsdc.FP64)
    r.read_sigma(path)
    assert r.time == 3600.0 and r.pdryini[0] == 0.0       # a modern header carries no pdryini
    r.advance(3)
    sr = r.get_state()
    # oracle from the same file
    o = oracle()
    st = sigio.model_state(head, data, o.sht)
    o.set_state(st["vrt"], st["div"], st["tv"], st["lnps"])
    o.set_topo(st["topo"])
    o.time = st["tstart"]
    o.advance(3)
    so = o.get_state()
    for k in ("vrt", "div", "tv", "lnps"):
        assert rel(sr[k][0], so[k]) < 1e-9, (k, rel(sr[k][0], so[k]))
        # FP32 file rounding against the uninterrupted run
        assert rel(sr[k][0], cont[k][0]) < 2e-5, (k, rel(sr[k][0], cont[k][0]))
    assert o.time == r.time == 3600.0 + 3 * 1200.0


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_dfi_matches_oracle(prec):
    import gfsdc
    p = gfsdc.FP64 if prec == "fp64" else gfsdc.FP32
    print(f"dfi {prec}", flush=True)
    o, g = oracle(), ctx(p)
    vrt, div, tv, lnps = cases.rest_state(o.sht.degree, o.sht.order, NL, seed=0, vort_rms=2e-5, nmax=10)
    rng = np.random.default_rng(5)
    sel = (o.sht.degree >= 8) & (o.sht.degree <= 20)
    lp = (rng.standard_normal(o.nlm) + 1j * rng.standard_normal(o.nlm)) * 3e-3 * sel
    lp[o.sht.order == 0] = lp[o.sht.order == 0].real
    o.set_state(vrt[0], div[0], tv[0], lnps[0] + lp)
    g.set_state(vrt, div, tv, lnps + lp[None])
    o.dfi(4)
    g.dfi(4, use_graph=True)
    assert o.time == g.time == 4 * 1200.0
    so, sg = o.get_state(), g.get_state()
    tol = 1e-9 if prec == "fp64" else 1e-5
    for k in ("vrt", "div", "tv", "lnps"):
        assert err(sg[k][0], so[k], prec) < tol, (k, err(sg[k][0], so[k], prec))
    # and the model goes on from the filtered state
    o.advance(2)
    g.advance(2, use_graph=True)
    so, sg = o.get_state(), g.get_state()
    for k in ("vrt", "div", "tv", "lnps"):
        assert err(sg[k][0], so[k], prec) < tol, (k, err(sg[k][0], so[k], prec))


@pytest.mark.parametrize("graph", [False, True])
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_iau_matches_oracle(prec, graph):
    """files at 0.5, 1, 1.5 h, dt 900 s from t = 0: the steps start before the
    window, on the first file, between files, on the interior file, on the last
    file and after the window"""
    import gfsdc
    p = gfsdc.FP64 if prec == "fp64" else gfsdc.FP32
    print(f"iau {prec} graph={graph}", flush=True)
    o, g = oracle(dt=900.0), ctx(p, dt=900.0)
    s = o.sht
    vrt, div, tv, lnps = cases.rest_state(s.degree, s.order, NL, seed=1, vort_rms=1e-5, nmax=15)
    rng = np.random.default_rng(8)

    def rnd(shape, scale):
        x = (rng.standard_normal(shape + (o.nlm,)) + 1j * rng.standard_normal(shape + (o.nlm,))) * scale
        x[..., s.order == 0] = x[..., s.order == 0].real
        x[..., s.degree > 20] = 0
        return x

    hrs = [0.5, 1.0, 1.5]
    inc = dict(vrt=rnd((3, NL), 2e-6), div=rnd((3, NL), 5e-7), tv=rnd((3, NL), 0.5), lnps=rnd((3,), 1e-3))
    o.set_state(vrt[0], div[0], tv[0], lnps[0])
    g.set_state(vrt, div, tv, lnps)
    o.iau_set(hrs, 6.0, inc["vrt"], inc["div"], inc["tv"], inc["lnps"])
    g.iau_set(hrs, 6.0, inc["vrt"][:, None], inc["div"][:, None], inc["tv"][:, None], inc["lnps"][:, None])
    ref = ctx(p, dt=900.0)
    ref.set_state(vrt, div, tv, lnps)
    o.advance(8)
    g.advance(8, use_graph=graph)
    ref.advance(8, use_graph=graph)
    so, sg, sr = o.get_state(), g.get_state(), ref.get_state()
    assert o.time == g.time == 8 * 900.0
    tol = 1e-9 if prec == "fp64" else 1e-5
    for k in ("vrt", "div", "tv", "lnps"):
        assert err(sg[k][0], so[k], prec) < tol, (k, err(sg[k][0], so[k], prec))
    # the increments did enter: the forced and unforced runs differ in T by
    # about the accumulated dt F (four steps' worth of a 0.5 K / 6 h increment)
    assert np.abs(sg["tv"] - sr["tv"]).max() > 1e-2
    # switched off, the context runs as one that never had it
    g.iau_set([], 6.0)
    g.set_state(vrt, div, tv, lnps)
    ref.set_state(vrt, div, tv, lnps)
    g.time = ref.time = 0.0
    g.advance(2, use_graph=graph)
    ref.advance(2, use_graph=graph)
    assert np.array_equal(g.get_state()["tv"], ref.get_state()["tv"])
