"""GPU parity of the shtns.f90 transform API (libgfsdc) against the FP64 oracle.

Tolerances: FP64 path 1e-11 relative (max-norm over the field batch);
FP32 path 2e-6 relative for scalar transforms and 5e-6 for vector ones
(2e-5 at T254, where the FP32 Legendre sums run over up to 256 degrees). The
FP32 vorticity/divergence analysis error is measured against the larger of
the two output fields: the wind is dominated by its rotational part, so FP32
rounding of u, v is of that size in both outputs.
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

GRIDS = [(128, 64, 42), (384, 190, 126), (768, 384, 254)]
A = 6.3712e6


def rand_spec(sht, nf, seed, slope=1.5):
    rng = np.random.default_rng(seed)
    n = np.maximum(sht.degree, 1).astype(float)
    q = n ** (-slope) * np.exp(2j * np.pi * rng.random((nf, sht.nlm)))
    q[:, sht.order == 0] = q[:, sht.order == 0].real
    return q


def rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


@pytest.fixture(scope="module", params=GRIDS, ids=lambda g: f"T{g[2]}")
def pair(request):
    import gfsdc
    nlon, nlat, T = request.param
    return O.Sht(nlon, nlat, T), gfsdc.Context(nlon, nlat, T, gfsdc.FP64), \
        gfsdc.Context(nlon, nlat, T, gfsdc.FP32)


def test_griddata(pair):
    o, g64, g32 = pair
    for g in (g64, g32):
        assert np.abs(g.lats - o.lats).max() < 1e-14
        assert np.abs(g.gauwts - o.gauwts).max() < 1e-14
        assert np.array_equal(g.degree, o.degree) and np.array_equal(g.order, o.order)


@pytest.mark.parametrize("nf", [1, 7, 70])
def test_spectogrd_grdtospec(pair, nf):
    o, g64, g32 = pair
    q = rand_spec(o, nf, 11 + nf)
    ref = o.spectogrd(q)
    got = g64.spectogrd(q)
    assert rel(got, ref) < 1e-11
    back = g64.grdtospec(got)
    assert rel(back, q) < 1e-11
    assert rel(g64.grdtospec(ref), o.grdtospec(ref)) < 1e-11
    got32 = g32.spectogrd(q.astype(np.complex64), dtype=1)
    assert rel(got32, ref) < 2e-6
    assert rel(g32.grdtospec(ref.astype(np.float32), dtype=1), o.grdtospec(ref)) < 2e-6


@pytest.mark.parametrize("nfw", [2, 4, 8])
@pytest.mark.parametrize("grid", GRIDS, ids=lambda g: f"T{g[2]}")
def test_fft_fields_per_workgroup(grid, nfw, monkeypatch):
    """Every FFT workgroup variant (2, 4 or 8 fields per workgroup: the tile
    autotuner's candidates, also set by GFSDC_FFT_NF) gives the oracle's
    transforms; 13 fields leave a partial last field group for each."""
    import gfsdc
    monkeypatch.setenv("GFSDC_FFT_NF", str(nfw))
    nlon, nlat, T = grid
    o = O.Sht(nlon, nlat, T)
    q = rand_spec(o, 13, 17)
    ref = o.spectogrd(q)
    back = o.grdtospec(ref)
    g64 = gfsdc.Context(nlon, nlat, T, gfsdc.FP64)
    assert rel(g64.spectogrd(q), ref) < 1e-11
    assert rel(g64.grdtospec(ref), back) < 1e-11
    g32 = gfsdc.Context(nlon, nlat, T, gfsdc.FP32)
    assert rel(g32.spectogrd(q.astype(np.complex64), dtype=1), ref) < 2e-6
    assert rel(g32.grdtospec(ref.astype(np.float32), dtype=1), back) < 2e-6


@pytest.mark.parametrize("nf", [1, 9])
def test_vector_transforms(pair, nf):
    o, g64, g32 = pair
    vrt = rand_spec(o, nf, 3) * 1e-5
    div = rand_spec(o, nf, 4) * 1e-6
    u0, v0 = o.getuv(vrt, div, A)
    u, v = g64.getuv(vrt, div, A)
    assert rel(u, u0) < 1e-11 and rel(v, v0) < 1e-11
    z0, d0 = o.getvrtdivspec(u0, v0, A)
    z, d = g64.getvrtdivspec(u0, v0, A)
    assert rel(z, z0) < 1e-11 and rel(d, d0) < 1e-11
    gx0, gy0 = o.getgrad(div, A)
    gx, gy = g64.getgrad(div, A)
    assert rel(gx, gx0) < 1e-11 and rel(gy, gy0) < 1e-11
    tol32 = 5e-6 if o.ntrunc < 200 else 2e-5  # FP32 sums over up to T+2 degrees
    u32, v32 = g32.getuv(vrt.astype(np.complex64), div.astype(np.complex64), A, dtype=1)
    assert rel(u32, u0) < tol32 and rel(v32, v0) < tol32
    z32, d32 = g32.getvrtdivspec(u0.astype(np.float32), v0.astype(np.float32), A, dtype=1)
    scale = max(np.abs(z0).max(), np.abs(d0).max())
    assert np.abs(z32 - z0).max() / scale < tol32 and np.abs(d32 - d0).max() / scale < tol32


@pytest.mark.parametrize("grid", GRIDS, ids=lambda g: f"T{g[2]}")
def test_bf16x3_transforms(grid):
    """BF16X3 precision: the Legendre GEMMs as three bf16 matrix-core products
    per FP32 product (hi.hi + hi.lo + lo.hi, lo.lo dropped, ~2^-16 relative
    per product). Tolerance (relative max-norm, T42 to T254): 3e-5 for scalar
    transforms; 1e-3 for getuv, whose extended-degree recurrence form sums
    terms that largely cancel (measured 2.6e-4 at T254), against 2e-5 for the
    FP32 path."""
    import gfsdc
    nlon, nlat, T = grid
    o = O.Sht(nlon, nlat, T)
    g = gfsdc.Context(nlon, nlat, T, gfsdc.BF16X3)
    q = rand_spec(o, 16, 5)
    ref = o.spectogrd(q)
    got = g.spectogrd(q.astype(np.complex64), dtype=1)
    e_syn = rel(got, ref)
    back = g.grdtospec(ref.astype(np.float32), dtype=1)
    e_ana = rel(back, o.grdtospec(ref))
    vrt, div = rand_spec(o, 4, 6), rand_spec(o, 4, 7) * 0.1
    vrt[:, 0] = div[:, 0] = 0
    u, v = g.getuv(vrt.astype(np.complex64), div.astype(np.complex64), A, dtype=1)
    uo, vo = o.getuv(vrt, div, A)
    e_uv = max(rel(u, uo), rel(v, vo))
    print(f"T{T} bf16x3 relative errors: synthesis {e_syn:.2e} analysis {e_ana:.2e} getuv {e_uv:.2e}")
    assert e_syn < 3e-5 and e_ana < 3e-5 and e_uv < 1e-3
    # and it is a genuinely different arithmetic path from FP32 (errors well above FP32 rounding)
    g32 = gfsdc.Context(nlon, nlat, T, gfsdc.FP32)
    assert rel(g32.spectogrd(q.astype(np.complex64), dtype=1), ref) < e_syn
