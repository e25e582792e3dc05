"""AR(1) spectral pattern of the oracle (patterngenerator.f90:33-177),
pinned by its statistics: unit-variance normalisation of the variance
spectrum, stationary grid variance amp^2, lag-1 autocorrelation phi =
exp(-dt/tau), spectrum shape exp(-L^2 n(n+1) / 4a^2), no global mean, and
the vertical taper of stoch_data.f90. The random stream is the counter RNG
(restated in numpy here), not the reference's Mersenne twister, so the
reference's individual draws are not reproduced: parity with the reference
is statistical only."""
import numpy as np

import cases
import levels
import oracle as O

NLON, NLAT, T, NL = 128, 64, 42, 20


def model(tau, amp=0.5, lscale=5.0e5, seed=3, dt=600.0):
    ak, bk = levels.hybrid_levels(NL)
    return O.Model(NLON, NLAT, T, NL, ak, bk, dt, tpat_amp=amp, tpat_tau=tau, tpat_lscale=lscale,
                   tpat_seed=seed)


def gridvar(m, spec):
    """computevarspec (patterngenerator.f90:70-89) with normfact = 2 pi"""
    w = np.where(m.sht.order == 0, 0.5, 1.0)
    return (w * np.abs(spec) ** 2).sum() / (2 * np.pi)


def test_varspectrum_normalised_and_shaped():
    m = model(21600.0)
    p = m.pattern()
    n = m.sht.degree.astype(float)
    a = m.cfg.rerth
    # shape: var^2 ratio between degrees follows exp(-L^2 n(n+1)/4a^2)
    r = p["varspectrum"] ** 2 / p["varspectrum"][0] ** 2
    np.testing.assert_allclose(r, np.exp(-(5.0e5) ** 2 * n * (n + 1) / (4 * a * a)), rtol=1e-12)
    # normalisation: the deterministic noise (1+i or sqrt 2)/sqrt(2n+1) has unit variance
    det = np.where(m.sht.order == 0, np.sqrt(2.0), 1 + 1j) / np.sqrt(2 * n + 1) * np.sqrt(2 * np.pi / T)
    det[0] = 0
    assert abs(gridvar(m, p["varspectrum"] * det) - 1.0) < 1e-12
    assert abs(p["phi"] - np.exp(-600.0 / 21600.0)) < 1e-15
    # vertical taper: 1 below sigma 0.1, 0 above sigma 0.05
    ak, bk = levels.hybrid_levels(NL)
    si = ak / 101300.0 + bk                       # interfaces top -> bottom (pressure_data.f90:64)
    sl = (0.5 * (si[:-1] + si[1:]))[::-1]         # layers bottom -> top
    vf = p["vfact"]
    assert np.all(vf[sl >= 0.1] == 1.0) and np.all(vf[sl <= 0.05] == 0.0)


def test_stationary_statistics():
    """short tau (phi = e^-1) for many independent samples"""
    m = model(600.0, amp=0.5)
    phi = m.pattern()["phi"]
    series = []
    for _ in range(1500):
        m.pattern_advance()
        series.append(m.pattern()["spec"].copy())
    S = np.array(series)
    assert np.all(S[:, 0] == 0)                                   # no global mean
    v = np.array([gridvar(m, s) for s in S])
    assert abs(v.mean() / 0.25 - 1.0) < 0.03                      # stationary variance amp^2
    # lag-1 autocorrelation of the coefficients
    x = np.concatenate([S.real[:, 1:], S.imag[:, m.sht.order != 0]], axis=1)
    c1 = (x[1:] * x[:-1]).sum() / (x * x).sum()
    assert abs(c1 - phi) < 0.02
    # variance spectrum by degree follows the varspectrum shape (sampling error allowed)
    w = np.where(m.sht.order == 0, 0.5, 1.0)
    n = m.sht.degree
    spec_n = np.array([(w[n == d] * (np.abs(S[:, n == d]) ** 2).mean(0)).sum() for d in range(T + 1)])
    shape = np.exp(-(5.0e5) ** 2 * np.arange(T + 1) * (np.arange(T + 1) + 1) / (4 * m.cfg.rerth ** 2))
    ratio = spec_n[1:] / shape[1:]
    assert np.std(ratio) / np.mean(ratio) < 0.1


def test_noise_stream_is_the_counter_rng():
    """the first draw of the pattern equals amp * varspectrum * noise built
    from the numpy restatement of the counter RNG (Box-Muller)"""
    m = model(21600.0, amp=1.0, seed=9)
    # rebuild init: noise(step 0) * amp * var, then nspinup advances
    p = m.pattern()
    n = m.sht.degree.astype(float)
    k = np.arange(m.nlm, dtype=np.uint64)

    def noise(step):
        u1 = cases.uniform(9, (1 << 40) + step, 2 * k)
        u2 = cases.uniform(9, (1 << 40) + step, 2 * k + 1)
        r = np.sqrt(-2 * np.log(1 - u1))
        z = np.where(m.sht.order != 0, r * np.cos(2 * np.pi * u2) + 1j * r * np.sin(2 * np.pi * u2),
                     np.sqrt(2.0) * r * np.cos(2 * np.pi * u2))
        z = z / np.sqrt(2 * n + 1) * np.sqrt(2 * np.pi / T)
        z[0] = 0
        return z

    q = 1.0 * p["varspectrum"] * noise(0)
    f = np.sqrt(1 - p["phi"] ** 2)
    for s in range(1, int(5 * 21600.0 / 600.0) + 1):
        q = p["phi"] * q + f * p["varspectrum"] * noise(s)
    assert np.abs(q - p["spec"]).max() < 1e-12 * np.abs(q).max()
