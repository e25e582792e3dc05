"""Digital filter initialisation and the incremental analysis update in the
oracle (run.f90:58-145 with set_dfi_wts :567-587; iau_module.f90:84-115 with
its calls at run.f90:250-258,302-309,363-370,429-436), pinned by:

1. set_dfi_wts against closed forms worked out by hand (N = 1, 2) and against
   the design of a Lanczos-windowed low-pass filter: unit gain at zero
   frequency, about half gain at the cut-off period 2 N dt, small gain beyond
   twice the cut-off frequency.
2. The DFI bookkeeping (weight k on the state after k steps, the clock at the
   window's middle) on a harmonic oscillator: for a sampled cosine
   x_k = cos(theta k + phi) and symmetric weights the filtered state is
   H(theta) x_N exactly, H the cosine transform of the weights.
3. DFI on the model removes gravity-wave noise: after a filter window over
   an unbalanced start, the surface pressure tendency is several times
   smaller than after the same number of plain steps.
4. The IAU schedule at hand-computed times (before the window, at the first
   file, between files, on an interior file, at the last file, after the
   window): one oracle step against the generic IMEX-RK step of imex_ref with
   the expected forcing added to every stage's tendency, and the clock.
"""
import numpy as np

import cases
import levels
import oracle as O
from imex_ref import imex_step, make


def H(w, theta):
    k = np.arange(w.size) - (w.size - 1) // 2
    return float((w * np.cos(theta * k)).sum())


def test_dfi_weights_closed_form():
    # N = 2: sx = pi (k-2)/2, wx = pi (k-2)/3; k = 1, 3: (2/pi)(3 sqrt3 / (2 pi)) = 3 sqrt3 / pi^2;
    # k = 0, 4: sin(pi) = 0 (to float32 pi); centre 1; normalised by 1 + 6 sqrt3 / pi^2
    w = O.dfi_weights(2)
    c = 1.0 / (1.0 + 6 * np.sqrt(3) / np.pi ** 2)
    np.testing.assert_allclose(w[[1, 2, 3]], [3 * np.sqrt(3) / np.pi ** 2 * c, c, 3 * np.sqrt(3) / np.pi ** 2 * c],
                               rtol=1e-6)
    assert abs(w[0]) < 1e-7 and abs(w[4]) < 1e-7
    # N = 1: sx = +-pi at both ends, so the filter is the identity
    np.testing.assert_allclose(O.dfi_weights(1), [0, 1, 0], atol=1e-7)


def test_dfi_weights_are_a_lanczos_low_pass():
    for N in (6, 12, 24):
        w = O.dfi_weights(N)
        assert abs(w.sum() - 1.0) < 1e-14 and np.allclose(w, w[::-1], rtol=0, atol=1e-17)
        thc = np.pi / N                                  # cut-off period 2 N dt
        assert abs(H(w, 0.0) - 1.0) < 1e-14
        assert 0.45 < H(w, thc) < 0.6
        assert 0.8 < H(w, 0.5 * thc) < 0.95
        assert max(abs(H(w, t)) for t in np.linspace(2 * thc, np.pi, 300)) < 0.05


class _Osc:
    """stand-in model: every state value a sampled cosine, cos(theta (t / dt) + phi)"""

    class _Cfg:
        dt = 600.0

    def __init__(self, theta, phi):
        self.cfg = self._Cfg()
        self.ntrac = 0
        self.theta, self.phi = theta, phi
        self.time = 0.0
        self.k = 0
        self.state = None

    def x(self, k):
        return np.cos(self.theta * k + self.phi)

    def get_state(self):
        if self.state is not None:
            return {k: v.copy() for k, v in self.state.items()}
        v = self.x(self.k)
        return dict(vrt=v, div=2 * v, tv=3 * v, trac=np.zeros(0), lnps=4 * v)

    def advance(self, n):
        self.k += n
        self.time += n * self.cfg.dt

    def set_state(self, vrt, div, tv, lnps, trac=None):
        self.state = dict(vrt=vrt, div=div, tv=tv, trac=np.zeros(0), lnps=lnps)


def test_dfi_bookkeeping_on_an_oscillator():
    N = 5
    w = O.dfi_weights(N)
    theta = np.array([0.05, 0.3, 0.9, 2.0])
    phi = np.array([0.1, -1.0, 2.5, 0.7])
    m = _Osc(theta, phi)
    m.time = 1000.0
    O.Model.dfi(m, N)
    want = np.array([H(w, t) for t in theta]) * np.cos(theta * N + phi)
    got = m.get_state()
    np.testing.assert_allclose(got["vrt"], want, rtol=0, atol=1e-14)
    np.testing.assert_allclose(got["lnps"], 4 * want, rtol=0, atol=1e-13)
    assert m.time == 1000.0 + N * 600.0 and m.k == 2 * N


def _unbalanced(nl):
    ak, bk = levels.hybrid_levels(nl)
    m = O.Model(64, 32, 21, nl, ak, bk, 1200.0, physics=1)
    s = m.sht
    vrt, div, tv, lnps = cases.rest_state(s.degree, s.order, nl, seed=0, vort_rms=2e-5, nmax=10)
    rng = np.random.default_rng(5)
    sel = (s.degree >= 8) & (s.degree <= 20)
    lp = (rng.standard_normal(s.nlm) + 1j * rng.standard_normal(s.nlm)) * 3e-3 * sel
    lp[s.order == 0] = lp[s.order == 0].real
    m.set_state(vrt[0], div[0], tv[0], lnps[0] + lp)
    return m


def test_dfi_removes_gravity_wave_noise():
    N = 8
    plain, filt = _unbalanced(10), _unbalanced(10)
    plain.advance(N)
    filt.dfi(N)
    assert filt.time == N * 1200.0
    plain.advance(1)
    filt.advance(1)
    p, f = plain.step_diag()["pstend"], filt.step_diag()["pstend"]
    assert p > 20.0 and f < p / 4, (p, f)


# IAU: files at 0.5, 1.0, 1.5 h, increments scaled by 1 / (6 h); the forcing at
# time t, by hand from getiauforcing, as weights on (inc0, inc1, inc2)
IAU_HRS = [0.5, 1.0, 1.5]
IAU_CASES = {0.0: None, 1800.0: (1, 0, 0), 2700.0: (0.5, 0.5, 0), 3600.0: (0, 1, 0),
             4500.0: (0, 0.5, 0.5), 5400.0: (0, 0, 1), 6300.0: None}


def test_iau_schedule_and_stage_forcing():
    nl = 8
    m = make(nl, 900.0)
    s = m.sht
    vrt, div, tv, lnps = cases.perturbed_state(s.degree, s.order, nl, seed=4, amp=1.0)
    y0 = (vrt[0], div[0], tv[0], lnps[0])
    rng = np.random.default_rng(9)

    def rnd(shape, scale):
        x = (rng.standard_normal(shape + (s.nlm,)) + 1j * rng.standard_normal(shape + (s.nlm,))) * scale
        x[..., s.order == 0] = x[..., s.order == 0].real
        return x

    inc = [dict(vrt=rnd((nl,), 2e-6), div=rnd((nl,), 1e-6), tv=rnd((nl,), 0.5), lnps=rnd((), 1e-3))
           for _ in IAU_HRS]
    m.iau_set(IAU_HRS, 6.0, [i["vrt"] for i in inc], [i["div"] for i in inc], [i["tv"] for i in inc],
              [i["lnps"] for i in inc])
    dti = 6.0 * 3600.0
    plain = make(nl, 900.0)
    for t, wts in IAU_CASES.items():
        if wts is None:
            want = imex_step(plain, y0)
        else:
            force = tuple(sum(c * i[k] for c, i in zip(wts, inc)) / dti for k in ("vrt", "div", "tv", "lnps"))
            want = imex_step(plain, y0, force=force)
        m.set_state(*y0)
        m.time = t
        m.advance(1)
        got = m.get_state()
        assert m.time == t + 900.0
        for name, w in zip(("vrt", "div", "tv", "lnps"), want):
            g, w = got[name][..., 1:], w[..., 1:]
            err = np.abs(g - w).max() / np.abs(w).max()
            assert err < 1e-11, (t, name, err)
    # the forcing moves the state by about dt F: the mid-window step against the plain one
    m.set_state(*y0)
    m.time = 2700.0
    m.advance(1)
    plain.set_state(*y0)
    plain.advance(1)
    dtf = 900.0 * 0.5 * (inc[0]["tv"] + inc[1]["tv"]) / dti
    diff = m.get_state()["tv"] - plain.get_state()["tv"]
    assert np.abs(diff - dtf).max() < 0.2 * np.abs(dtf).max()


def test_iau_rejects_uneven_or_single_files():
    m = make(4, 900.0)
    z = np.zeros((2, 4, m.nlm), np.complex128)
    for hrs in ([0.0, 1.0, 3.0], [3.0]):
        n = len(hrs)
        zz = np.zeros((n, 4, m.nlm), np.complex128)
        try:
            m.iau_set(hrs, 6.0, zz, zz, zz, np.zeros((n, m.nlm), np.complex128))
        except ValueError:
            continue
        raise AssertionError(f"accepted {hrs}")
    m.iau_set([0.0, 1.0], 6.0, z, z, z, np.zeros((2, m.nlm), np.complex128))
    m.iau_set([], 6.0, None, None, None, None)
