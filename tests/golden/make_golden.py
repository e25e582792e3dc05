"""Writes the golden fixtures used by tests/test_oracle_*.py.

Sources are independent of this repository's code: Gauss-Legendre nodes and
weights by Newton iteration on P_nlat(x) in mpmath at 40 digits (started from
scipy.special.roots_legendre, whose own polar weights are only good to ~4e-10
relative at nlat = 384), and for the associated Legendre functions
the explicit sum of the m-th derivative of the Legendre polynomial,
P_n(x) = 2^-n sum_k (-1)^k C(n,k) C(2n-2k,n) x^(n-2k), evaluated in mpmath at
600 digits (the alternating sum cancels ~150 digits at n = 255), normalised
to the orthonormal, no-Condon-Shortley-phase convention that src/shtns.f90
requests from SHTns (shtns.f90:110):
Pbar_n^m(x) = sqrt((2n+1)/(4 pi) (n-m)!/(n+m)!) (1-x^2)^(m/2) d^m P_n/dx^m.

Run: python tests/golden/make_golden.py   (writes tests/golden/*.npz)
"""
import os

import mpmath as mp
import numpy as np
from scipy.special import roots_legendre

HERE = os.path.dirname(os.path.abspath(__file__))


def gauss():
    mp.mp.dps = 40
    out = {}
    for nlat in (64, 190, 384):
        x0, _ = roots_legendre(nlat)
        xs, ws = [], []
        for xi in np.sort(x0)[::-1]:
            x = mp.mpf(xi)
            for _ in range(8):
                p0, p1 = mp.mpf(1), x
                for k in range(2, nlat + 1):
                    p0, p1 = p1, ((2 * k - 1) * x * p1 - (k - 1) * p0) / k
                dp = nlat * (x * p1 - p0) / (x * x - 1)
                x = x - p1 / dp
            p0, p1 = mp.mpf(1), x
            for k in range(2, nlat + 1):
                p0, p1 = p1, ((2 * k - 1) * x * p1 - (k - 1) * p0) / k
            dp = nlat * (x * p1 - p0) / (x * x - 1)
            xs.append(float(x))
            ws.append(float(2 / ((1 - x * x) * dp * dp)))
        out[f"mu_{nlat}"] = np.array(xs)
        out[f"w_{nlat}"] = np.array(ws)
    np.savez(os.path.join(HERE, "gauss_nodes.npz"), **out)


def pbar(n, m, x):
    norm = mp.sqrt((2 * n + 1) / (4 * mp.pi) * mp.factorial(n - m) / mp.factorial(n + m))
    d = mp.mpf(0)
    for k in range(n // 2 + 1):
        p = n - 2 * k
        if p < m:
            break
        d += (-1) ** k * mp.binomial(n, k) * mp.binomial(2 * n - 2 * k, n) * \
            mp.factorial(p) / mp.factorial(p - m) * x ** (p - m)
    return norm * (1 - x * x) ** (mp.mpf(m) / 2) * d / mp.mpf(2) ** n


def legendre():
    mp.mp.dps = 600
    T = 254
    x, _ = roots_legendre(384)
    mu = np.sort(x)[::-1]
    js = [0, 1, 10, 60, 191]
    ms = [0, 1, 17, 128, 200, 254]
    rows = []
    for j in js:
        xm = mp.mpf(mu[j])
        for m in ms:
            for n in range(m, T + 2):
                v = pbar(n, m, xm)
                rows.append((j, m, n, float(v) if abs(v) > mp.mpf("1e-300") else 0.0))
    a = np.array(rows, dtype=[("j", "i4"), ("m", "i4"), ("n", "i4"), ("p", "f8")])
    np.savez(os.path.join(HERE, "legendre_T254.npz"), table=a, mu=mu)


if __name__ == "__main__":
    import sys
    if "--gauss-only" in sys.argv:
        gauss()
    else:
        gauss()
        legendre()
