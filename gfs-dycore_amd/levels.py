"""Vertical coordinates and synthetic initial states for the benchmark configs.

The reference reads ak/bk from its initial-condition file header
(dyn_init.f90:52-61, reversed to top -> bottom) and those files are not
distributed, so the configs define their own coordinates:

* ``sigma_levels(n)``: n equally spaced sigma layers (ak = 0, bk = 0..1), the
  Held-Suarez (1994) layout used by the T42L20 configuration.
* ``hybrid_levels(n)``: n layers equally spaced in eta with a hybrid split
  B = ((eta - 0.2) / 0.8) ** 1.5 below eta = 0.2 and A = p0 (eta - B), so the
  top interface is p = 0 (as GFS), the upper fifth of the column is pure
  pressure and the surface is pure sigma. Used by the L64 configurations.

Both return (ak, bk) at the nlev+1 interfaces, top -> bottom, ak in Pa.
"""
import numpy as np


def sigma_levels(nlev):
    ak = np.zeros(nlev + 1)
    bk = np.linspace(0.0, 1.0, nlev + 1)
    return ak, bk


def hybrid_levels(nlev, eta_c=0.2, expo=1.5, p0=1.0e5):
    eta = np.linspace(0.0, 1.0, nlev + 1)
    bk = np.clip((eta - eta_c) / (1.0 - eta_c), 0.0, None) ** expo
    bk[-1] = 1.0
    ak = p0 * (eta - bk)
    ak[-1] = 0.0
    return ak, bk


# grid and time step of the configurations (dt in seconds)
CONFIGS = {
    "T42L20": dict(nlon=128, nlat=64, ntrunc=42, nlev=20, dt=1200.0, levels="sigma"),
    "T126L64": dict(nlon=384, nlat=190, ntrunc=126, nlev=64, dt=600.0, levels="hybrid"),
    "T254L64": dict(nlon=768, nlat=384, ntrunc=254, nlev=64, dt=600.0, levels="hybrid"),
}


def moist_grid(lats, nlon, ak, bk, q0=0.025, seed=0):
    """Specific humidity on the grid, shape (nlev bottom -> top, nlat, nlon): a
    moist tropical lower troposphere drying upward (sigma^3) and poleward, with a
    zonal wavenumber-3 modulation whose phase depends on the seed; supersaturated
    near the surface in the tropics of a 300 K state, so condensation occurs."""
    ak = np.asarray(ak, np.float64)
    bk = np.asarray(bk, np.float64)
    sig = (0.5 * ((ak[:-1] + ak[1:]) / 1.0e5 + (bk[:-1] + bk[1:])))[::-1]
    ph = 2 * np.pi * _uniform(seed, 7, 0)
    lam = 2 * np.pi * np.arange(nlon) / nlon
    lat = np.asarray(lats, np.float64)
    return (q0 * np.exp(-(lat / 0.7) ** 4)[None, :, None] * sig[:, None, None] ** 3 *
            (1.0 + 0.2 * np.cos(3 * lam + ph))[None, None, :])


def _uniform(seed, stream, index):
    """one draw of the library's counter RNG (splitmix64 of seed, stream, index)"""
    m = (1 << 64) - 1
    z = ((seed * 0x9E3779B97F4A7C15) & m) ^ (((stream + 0x632BE59BD9B4E019) & m) * 0xBF58476D1CE4E5B9 & m)
    z = (z + index * 0x94D049BB133111EB + 0x9E3779B97F4A7C15) & m
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & m
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & m
    z ^= z >> 31
    return (z >> 11) * 2.0 ** -53


def levels_for(name):
    c = CONFIGS[name]
    return sigma_levels(c["nlev"]) if c["levels"] == "sigma" else hybrid_levels(c["nlev"])
