"""FP64 column MLP of the oracle (oracle/mlp.c), the checker of the GPU
network's forward, backward and AdamW kernels. No reference counterpart;
pinned by central finite differences of the loss, a hand-computed
single-layer case and the AdamW closed form."""
import numpy as np

import oracle as O


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    net = O.Mlp(7, 4, 6, 2)
    p = rng.normal(0, 0.5, net.nparams)
    x = rng.normal(0, 1, (5, 7))
    t = rng.normal(0, 1, (5, 4))
    loss, g, _ = net.loss_grad(p, x, t)
    h = 1e-6
    idx = rng.choice(net.nparams, 60, replace=False)
    fd = np.array([(net.loss_grad(p + h * np.eye(1, net.nparams, i)[0], x, t, False)[0] -
                    net.loss_grad(p - h * np.eye(1, net.nparams, i)[0], x, t, False)[0]) / (2 * h) for i in idx])
    np.testing.assert_allclose(g[idx], fd, rtol=1e-6, atol=1e-9)


def test_linear_network_closed_form():
    """no hidden layer (nlayers = 0): y = x W + b, loss and gradient by hand"""
    rng = np.random.default_rng(1)
    net = O.Mlp(3, 2, 32, 0)
    W = rng.normal(size=(3, 2))
    b = rng.normal(size=2)
    p = np.concatenate([W.ravel(), b])
    x = rng.normal(size=(4, 3))
    t = rng.normal(size=(4, 2))
    loss, g, y = net.loss_grad(p, x, t)
    yr = x @ W + b
    np.testing.assert_allclose(y, yr, rtol=1e-14)
    assert abs(loss - ((yr - t) ** 2).mean()) < 1e-14
    d = 2 * (yr - t) / yr.size
    np.testing.assert_allclose(g, np.concatenate([(x.T @ d).ravel(), d.sum(0)]), rtol=1e-13)


def test_gelu_network_small_case():
    """one GELU hidden layer evaluated in numpy (erf GELU)"""
    from math import erf
    rng = np.random.default_rng(2)
    net = O.Mlp(3, 2, 4, 1)
    W1, b1, W2, b2 = rng.normal(size=(3, 4)), rng.normal(size=4), rng.normal(size=(4, 2)), rng.normal(size=2)
    p = np.concatenate([W1.ravel(), b1, W2.ravel(), b2])
    x = rng.normal(size=(3, 3))
    gelu = np.vectorize(lambda z: 0.5 * z * (1 + erf(z / np.sqrt(2))))
    yr = gelu(x @ W1 + b1) @ W2 + b2
    _, _, y = net.loss_grad(p, x, np.zeros((3, 2)), False)
    np.testing.assert_allclose(y, yr, rtol=1e-13)


def test_adamw_closed_form():
    rng = np.random.default_rng(3)
    net = O.Mlp(3, 2, 4, 1)
    p0 = rng.normal(size=net.nparams)
    p, m, v = p0.copy(), np.zeros(net.nparams), np.zeros(net.nparams)
    g1, g2 = rng.normal(size=net.nparams), rng.normal(size=net.nparams)
    lr, wd = 1e-2, 0.1
    net.adamw(p, m, v, g1, 1, lr, wd)
    net.adamw(p, m, v, g2, 2, lr, wd)
    # by hand: weights decay, biases do not
    isw = np.concatenate([np.ones(12), np.zeros(4), np.ones(8), np.zeros(2)]).astype(bool)
    q, mm, vv = p0.copy(), np.zeros_like(p0), np.zeros_like(p0)
    for s, g in ((1, g1), (2, g2)):
        mm = 0.9 * mm + 0.1 * g
        vv = 0.999 * vv + 0.001 * g * g
        q = np.where(isw, q * (1 - lr * wd), q)
        q = q - lr * (mm / (1 - 0.9 ** s)) / (np.sqrt(vv / (1 - 0.999 ** s)) + 1e-8)
    np.testing.assert_allclose(p, q, rtol=1e-14, atol=1e-15)


def _bf16(a):
    """round to bfloat16 (nearest even) through float32"""
    u = np.ascontiguousarray(np.asarray(a, np.float64).astype(np.float32)).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def test_bf16_storage_mode():
    """the oracle's bf16 mode rounds exactly where a bf16 matrix-core network
    stores values: inputs, weights, biases, hidden activations after GELU, the
    output gradient and each back-propagated pre-activation gradient (written
    out here in numpy over two hidden layers, forward and backward)"""
    from scipy.special import erf
    rng = np.random.default_rng(4)
    nin, nout, hid, nlay, n = 9, 5, 8, 2, 6
    net = O.Mlp(nin, nout, hid, nlay)
    p = rng.normal(0, 0.4, net.nparams)
    x = rng.normal(0, 1, (n, nin))
    t = rng.normal(0, 1, (n, nout))
    loss, g, y = net.loss_grad(p, x, t, bf16=True)
    dims = [(nin, hid), (hid, hid), (hid, nout)]
    Ws, bs, o = [], [], 0
    for K, N in dims:
        Ws.append(_bf16(p[o:o + K * N].reshape(K, N)))
        bs.append(_bf16(p[o + K * N:o + K * N + N]))
        o += K * N + N
    gelu = lambda z: 0.5 * z * (1 + erf(z / np.sqrt(2)))
    dgelu = lambda z: 0.5 * (1 + erf(z / np.sqrt(2))) + z * np.exp(-0.5 * z * z) / np.sqrt(2 * np.pi)
    acts, pres = [_bf16(x)], []
    for l, (W, b) in enumerate(zip(Ws, bs)):
        z = acts[-1] @ W + b
        pres.append(z)
        acts.append(_bf16(gelu(z)) if l < len(Ws) - 1 else z)
    np.testing.assert_allclose(y, acts[-1], rtol=1e-12, atol=1e-14)
    assert abs(loss - ((acts[-1] - t) ** 2).mean()) < 1e-12
    d = _bf16(2 * (acts[-1] - t) / (n * nout))
    gr = []
    for l in range(len(Ws) - 1, -1, -1):
        gr = [(acts[l].T @ d).ravel(), d.sum(0)] + gr
        if l > 0:
            d = _bf16((d @ Ws[l].T) * dgelu(pres[l - 1]))
    np.testing.assert_allclose(g, np.concatenate(gr), rtol=1e-11, atol=1e-15)
    # and the rounding is real: the plain FP64 network differs at the bf16 level
    _, _, y64 = net.loss_grad(p, x, t, grad=False)
    assert 1e-4 < np.abs(y - y64).max() / np.abs(y64).max() < 5e-2
