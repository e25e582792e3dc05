"""Shared test inputs: seeded spectral states built in numpy (independent of
both the product and the oracle) and the counter RNG restated in numpy."""
import numpy as np

M64 = (1 << 64) - 1


def splitmix(seed, stream, index):
    """numpy restatement of the counter RNG (splitmix64 finaliser), uint64."""
    with np.errstate(over="ignore"):
        s = np.uint64(seed)
        z = (s * np.uint64(0x9E3779B97F4A7C15)) ^ \
            ((np.asarray(stream, np.uint64) + np.uint64(0x632BE59BD9B4E019)) *
             np.uint64(0xBF58476D1CE4E5B9))
        z = z + np.asarray(index, np.uint64) * np.uint64(0x94D049BB133111EB) + \
            np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def uniform(seed, stream, index):
    return (splitmix(seed, stream, index) >> np.uint64(11)).astype(np.float64) * 2.0 ** -53


def rest_state(degree, order, nlev, t0=300.0, ps0=1.0e5, vort_rms=1e-6, nmin=1, nmax=21,
               seed=0, nmem=1):
    """Resting isothermal state + random vorticity in total wavenumbers
    [nmin, nmax], rms grid vorticity vort_rms per level (orthonormal norm:
    mean of f^2 over the sphere = (sum_{m=0} |q|^2 + 2 sum_{m>0} |q|^2) / 4 pi)."""
    nlm = degree.size
    sq = np.sqrt(4 * np.pi)
    vrt = np.zeros((nmem, nlev, nlm), np.complex128)
    k = np.arange(nlm)
    sel = (degree >= nmin) & (degree <= nmax)
    for mem in range(nmem):
        for lev in range(nlev):
            th = 2 * np.pi * uniform(seed + mem, lev, k)
            x = np.where(order == 0, np.cos(th) + 0j, np.exp(1j * th))
            x[~sel] = 0
            w = np.where(order == 0, 1.0, 2.0)
            rms = np.sqrt((w * np.abs(x) ** 2).sum() / (4 * np.pi))
            vrt[mem, lev] = x * (vort_rms / rms) if rms > 0 else 0.0
    div = np.zeros_like(vrt)
    tv = np.zeros_like(vrt)
    tv[:, :, 0] = t0 * sq
    lnps = np.zeros((nmem, nlm), np.complex128)
    lnps[:, 0] = np.log(ps0) * sq
    return vrt, div, tv, lnps


def perturbed_state(degree, order, nlev, seed, amp=1.0, nmax=None):
    """A rough, fully three-dimensional state: rest state plus random
    vorticity, divergence, temperature and ln(ps) perturbations."""
    rng = np.random.default_rng(seed)
    nlm = degree.size
    nmax = degree.max() if nmax is None else nmax

    def rnd(scale, shape):
        n = np.maximum(degree, 1).astype(float)
        x = rng.standard_normal(shape + (nlm,)) + 1j * rng.standard_normal(shape + (nlm,))
        x = x * scale * n ** -1.0
        x[..., order == 0] = x[..., order == 0].real
        x[..., degree > nmax] = 0
        x[..., 0] = 0
        return x

    vrt, div, tv, lnps = rest_state(degree, order, nlev, vort_rms=0.0, nmin=1, nmax=0)
    vrt = vrt + amp * rnd(2e-5, (1, nlev))
    div = div + amp * rnd(2e-6, (1, nlev))
    tv = tv + amp * rnd(3.0, (1, nlev))
    lnps = lnps + amp * rnd(3e-3, (1,))
    return vrt, div, tv, lnps


def moist_grid(lats, nlon, ak, bk, q0=0.025, seed=0):
    """Specific humidity on the grid (nlev bottom -> top, nlat, nlon); the
    synthetic moist state of the benchmark configurations (levels.moist_grid)."""
    import levels
    return levels.moist_grid(lats, nlon, ak, bk, q0=q0, seed=seed)
