"""Learned column physics on the GPU: the bf16 MFMA forward, backward and
AdamW kernels against the FP64 oracle network, training on columns the plain
simple physics generates, the gain over the method without the network, and
the judge / fallback inside advance().

Tolerances: the GPU network stores inputs, weights, hidden activations and
back-propagated gradients in bf16 (8-bit mantissa) and sums in FP32. The
oracle network is evaluated with the same bf16 storage rounding (sums in
FP64), so what remains is FP32 summation and the rare activation that the
GPU's fast GELU rounds to the neighbouring bf16 value: 4e-3 relative L2
(forward) and 6e-3 (gradients). AdamW runs in FP32 and is compared with its
closed form at 1e-5.
"""
import numpy as np
import pytest

import cases
import levels
import oracle as O

pytestmark = pytest.mark.gpu

NLON, NLAT, T, NL, DT = 128, 64, 42, 20, 1200.0
HID, NLAY = 64, 2


def bf16(a):
    """round to bfloat16 (nearest even) and back, as the GPU stores weights"""
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def rel2(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def moist_ctx(prec=None, seed=3, **kw):
    import gfsdc
    ak, bk = levels.hybrid_levels(NL)
    g = gfsdc.Context(NLON, NLAT, T, gfsdc.FP32 if prec is None else prec, nlev=NL, ak=ak, bk=bk, dt=DT,
                      ntrac=1, physics=gfsdc.PHYS_SIMPLE, simple_test=2, turbflux=1, **kw)
    sht = O.Sht(NLON, NLAT, T)
    vrt, div, tv, lnps = cases.perturbed_state(sht.degree, sht.order, NL, seed=seed, amp=0.5)
    q = sht.grdtospec(cases.moist_grid(sht.lats, NLON, ak, bk, seed=seed))
    g.set_state(vrt, div, tv, lnps, trac=q[None, None])
    return g


def layers(p, nin, nout):
    out, o = [], 0
    for K, N in [(nin, HID)] + [(HID, HID)] * (NLAY - 1) + [(HID, nout)]:
        out.append(p[o:o + K * N + N])
        o += K * N + N
    return out


def test_forward_matches_oracle():
    g = moist_ctx()
    g.mlp_init(HID, NLAY, 256, 1000, seed=5)
    nin, nout = g.mlp_nin, g.mlp_nout
    rng = np.random.default_rng(0)
    x = rng.normal(0, 1, (300, nin)).astype(np.float32)     # more rows than one batch
    y = g.mlp_forward(x)
    net = O.Mlp(nin, nout, HID, NLAY)
    _, _, yo = net.loss_grad(g.mlp_params(), x, np.zeros((300, nout)), grad=False, bf16=True)
    print(f"forward rel L2 vs bf16 oracle: {rel2(y, yo):.2e}")
    assert rel2(y, yo) < 4e-3, rel2(y, yo)
    # a network that skipped the rounding, or mis-scaled a layer, is far off
    _, _, y64 = net.loss_grad(g.mlp_params(), x, np.zeros((300, nout)), grad=False)
    assert rel2(y64, yo) > 2 * rel2(y, yo)


def test_gradient_matches_oracle():
    g = moist_ctx()
    g.mlp_init(HID, NLAY, 512, 1000, seed=6)
    nin, nout = g.mlp_nin, g.mlp_nout
    rng = np.random.default_rng(1)
    x = rng.normal(0, 1, (400, nin)).astype(np.float32)
    t = rng.normal(0, 1, (400, nout)).astype(np.float32)
    grad, loss = g.mlp_grad(x, t)
    net = O.Mlp(nin, nout, HID, NLAY)
    lo, go, _ = net.loss_grad(g.mlp_params(), x, t, bf16=True)
    assert abs(loss - lo) < 1e-3 * lo, (loss, lo)
    for a, b in zip(layers(grad, nin, nout), layers(go, nin, nout)):
        print(f"gradient block rel L2 vs bf16 oracle: {rel2(a, b):.2e}")
        assert rel2(a, b) < 6e-3, rel2(a, b)


def test_adamw_matches_closed_form():
    g = moist_ctx()
    g.mlp_init(HID, NLAY, 256, 1000, seed=7)
    nin, nout = g.mlp_nin, g.mlp_nout
    rng = np.random.default_rng(2)
    p0 = g.mlp_params(set=rng.normal(0, 0.1, g.mlp_params().size).astype(np.float32)).astype(np.float64)
    g1 = rng.normal(0, 1e-3, p0.size)
    g2 = rng.normal(0, 1e-3, p0.size)
    lr, wd = 1e-3, 0.05
    g.mlp_adamw(g1, lr, wd)
    g.mlp_adamw(g2, lr, wd)
    p = g.mlp_params().astype(np.float64)
    net = O.Mlp(nin, nout, HID, NLAY)
    q, m, v = p0.copy(), np.zeros_like(p0), np.zeros_like(p0)
    net.adamw(q, m, v, g1.astype(np.float32).astype(np.float64), 1, lr, wd)
    net.adamw(q, m, v, g2.astype(np.float32).astype(np.float64), 2, lr, wd)
    assert np.abs(p - q).max() < 1e-5 * max(np.abs(q).max(), 1.0)


def test_trained_network_beats_the_method_without_it_and_judges():
    """Train on columns of the plain T42L20 moist model (after a 4-hour spin-up
    that removes the initial saturation adjustment). The network's normalised
    rms error is below that of the method without it (the mean physics change)
    on the training columns (by half) and on held-out columns from the steps
    that follow them (by 15 %); inside advance() the judge measures it."""
    g = moist_ctx()
    g.advance(12)
    g.mlp_init(128, 2, 1024, 200000, seed=11)
    n = g.mlp_collect(48, 4)                      # 2048 columns per step
    g.mlp_normalize()
    ntrain = int(n * 0.85)
    losses = g.mlp_train(3000, 1024, 1e-3, 1e-4, ntrain=ntrain)
    assert losses[-100:].mean() < 0.5 * losses[:20].mean()
    err_tr, base_tr = g.mlp_eval(0, 20000)
    err, base = g.mlp_eval(ntrain, n - ntrain)    # later in time than any training column
    print(f"normalised rms error: train {err_tr:.3f} (mean predictor {base_tr:.3f}), "
          f"held out {err:.3f} (mean predictor {base:.3f})")
    assert err_tr < 0.6 * base_tr, (err_tr, base_tr)
    assert err < 0.85 * base, (err, base)
    g.mlp_enable(True, judge_stride=16, threshold=10.0)
    g.advance(3, use_graph=True)
    st = g.mlp_judge_stats()
    assert st["judged"] == 3 and st["fallbacks"] == 0 and np.isfinite(st["error"])
    s = g.get_state()
    assert all(np.isfinite(s[k]).all() for k in ("vrt", "div", "tv", "lnps"))


def test_fallback_reproduces_the_plain_physics():
    """threshold 0: every judged step falls back, and advance() then matches a
    context that never used the network, bit for bit"""
    a, b = moist_ctx(seed=4), moist_ctx(seed=4)
    a.mlp_init(HID, NLAY, 512, 20000, seed=3)
    a.mlp_collect(2, 8)
    a.mlp_normalize()
    # both contexts restart from the same state after the collection steps
    st = a.get_state()
    a.set_state(st["vrt"], st["div"], st["tv"], st["lnps"], trac=st["trac"])
    b.set_state(st["vrt"], st["div"], st["tv"], st["lnps"], trac=st["trac"])
    a.pdryini = b.pdryini = 0.0
    a.mlp_enable(True, judge_stride=8, threshold=0.0)
    a.advance(2)
    b.advance(2)
    sa, sb = a.get_state(), b.get_state()
    assert a.mlp_judge_stats()["fallbacks"] == 2
    for k in ("vrt", "div", "tv", "lnps", "trac"):
        assert np.array_equal(sa[k], sb[k]), k


def test_network_moves_to_another_context():
    """a network trained in one context, installed in another through
    gfsdc_mlp_params and gfsdc_mlp_norm, computes the same outputs and the same
    judged steps; the normalisation refuses non-positive spreads"""
    import gfsdc
    print("mlp transfer", flush=True)
    a = moist_ctx(seed=6)
    a.mlp_init(HID, NLAY, 1024, 20000, seed=5)
    a.mlp_collect(6, 8)
    a.mlp_normalize()
    a.mlp_train(200, 256, 1e-3, 1e-4)
    p, nrm = a.mlp_params(), a.mlp_norm()
    assert nrm.size == 2 * (a.mlp_nin + a.mlp_nout)
    b = moist_ctx(seed=6)
    b.mlp_init(HID, NLAY, 1024, 1, seed=99)       # other initial weights, no data
    with pytest.raises(gfsdc.GfsdcError, match="normalisation"):
        b.mlp_enable(True)                         # nothing installed yet
    b.mlp_params(set=p)
    bad = nrm.copy()
    bad[a.mlp_nin] = 0.0                           # a zero input spread
    with pytest.raises(gfsdc.GfsdcError, match="positive spread"):
        b.mlp_norm(set=bad)
    b.mlp_norm(set=nrm)
    assert np.array_equal(b.mlp_norm(), nrm)
    assert np.array_equal(b.mlp_params(), p)
    x = np.random.default_rng(2).standard_normal((300, a.mlp_nin)).astype(np.float32)
    assert np.array_equal(a.mlp_forward(x), b.mlp_forward(x))
    # the same judged steps from the same state
    c = moist_ctx(seed=7)
    st = c.get_state()
    for g in (a, b):
        g.set_state(st["vrt"], st["div"], st["tv"], st["lnps"], trac=st["trac"])
        g.pdryini = np.zeros(1)                    # both fix the dry mass of this state
        g.mlp_enable(True, judge_stride=16, threshold=10.0)
        g.advance(2, use_graph=True)
    sa, sb = a.get_state(), b.get_state()
    for k in ("vrt", "div", "tv", "lnps", "trac"):
        assert np.array_equal(sa[k], sb[k]), k
    # (the judged columns differ: each context keeps its own sampling counter)
    assert a.mlp_judge_stats()["judged"] == b.mlp_judge_stats()["judged"] == 2
