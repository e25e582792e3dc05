"""gfsdc_spectogrd_noalloc -- the synthesis for callers that cannot allocate
device memory -- against gfsdc_spectogrd, the entry point whose kernels it
shares, through the C ABI: bit-identical grids on two calls in a row, the
second with fewer fields (its scratch keeps the first call's contents), in
FP64 at T126 and FP32 at T254; the FP64 grids also equal the oracle to 1e-11;
a call that needs more than the scratch holds is refused before any work.

The caller's device buffers come from the HIP runtime that libgfsdc links
(libamdhip64.so.7), not from torch, whose wheel carries a runtime of its own."""
import ctypes as C

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def lib():
    import gfsdc
    L = gfsdc.load()
    L.gfsdc_spectogrd_noalloc.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]
    return L


def rand_spec(degree, order, nf, seed):
    rng = np.random.default_rng(seed)
    n = np.maximum(degree, 1).astype(float)
    q = n ** -1.5 * np.exp(2j * np.pi * rng.random((nf, degree.size)))
    q[:, order == 0] = q[:, order == 0].real
    return q


def hip():
    import gfsdc
    gfsdc.load()                        # libgfsdc first: its runtime is the one loaded
    H = C.CDLL("libamdhip64.so.7")
    H.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    H.hipFree.argtypes = [C.c_void_p]
    H.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return H


def noalloc(ctx, q, real):
    """q: (nf, nlm) complex spectra of the context's element type; returns the grids"""
    H, L = hip(), lib()
    H2D, D2H = 1, 2
    q = np.ascontiguousarray(q)
    out = np.full((q.shape[0], ctx.nlat, ctx.nlon), np.nan, real)
    ds, dg = C.c_void_p(), C.c_void_p()
    assert H.hipMalloc(C.byref(ds), q.nbytes) == 0 and H.hipMalloc(C.byref(dg), out.nbytes) == 0
    try:
        assert H.hipMemcpy(ds, q.ctypes.data, q.nbytes, H2D) == 0
        assert H.hipMemcpy(dg, out.ctypes.data, out.nbytes, H2D) == 0      # NaN where nothing is written
        rc = L.gfsdc_spectogrd_noalloc(ctx.h, ds, dg, q.shape[0])
        assert rc == 0, L.gfsdc_last_error()
        ctx.sync()
        assert H.hipMemcpy(out.ctypes.data, dg, out.nbytes, D2H) == 0
    finally:
        H.hipFree(ds)
        H.hipFree(dg)
    return out


@pytest.mark.parametrize("prec,shape,nfs", [("fp64", (384, 190, 126), (24, 5)),
                                            ("fp32", (768, 384, 254), (16, 3))])
def test_noalloc_synthesis_equals_spectogrd(prec, shape, nfs):
    import gfsdc
    nlon, nlat, T = shape
    p = gfsdc.FP64 if prec == "fp64" else gfsdc.FP32
    real = np.float64 if prec == "fp64" else np.float32
    ctx = gfsdc.Context(nlon, nlat, T, p)
    o = O.Sht(nlon, nlat, T) if prec == "fp64" else None
    for call, nf in enumerate(nfs):
        print(f"noalloc {prec} T{T}: call {call} with {nf} fields", flush=True)
        q = rand_spec(ctx.degree, ctx.order, nf, seed=10 + call)
        if prec == "fp32":
            q = q.astype(np.complex64)
        got = noalloc(ctx, q, real)
        want = ctx.spectogrd(q, dtype=gfsdc.DT_F64 if prec == "fp64" else gfsdc.DT_F32)
        assert got.shape == want.shape and np.array_equal(got, want), (call, np.nanmax(np.abs(got - want)))
        if o is not None:
            ref = o.spectogrd(q)
            assert np.abs(got - ref).max() < 1e-11 * np.abs(ref).max()


def test_noalloc_refuses_oversize_requests():
    import gfsdc
    L = lib()
    ctx = gfsdc.Context(768, 384, 254, gfsdc.FP32)
    dummy = C.c_void_p(16)   # never dereferenced: the size check comes first
    rc = L.gfsdc_spectogrd_noalloc(ctx.h, dummy, dummy, 400)
    assert rc == -1                           # GFSDC_EINVAL
    assert b"MiB of scratch" in L.gfsdc_last_error()
