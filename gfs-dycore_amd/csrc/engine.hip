// Engine<T>: context set-up, caller-buffer staging and the batched per-field
// transforms of the shtns.f90 API (spectogrd, grdtospec, getuv,
// getvrtdivspec, getgrad) on the device.
#include <cstring>
#include <mutex>

#include "engine.h"

namespace gfsdc {

// ---------------------------------------------------------------------------
// Scratch of the no-allocation synthesis (gfsdc_spectogrd_noalloc): one
// file-scope device array per GPU, held until the process ends, holding the
// synthesis descriptors, the extended spectral columns and the Fourier buffer
// that gfsdc_spectogrd keeps in the context's lazily allocated work buffers.
// Its device address is looked up once per device at context creation, with
// that device current; host code uses only that address, never the symbol.
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];
constexpr int MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static void *g_scratch_ptr[MAX_DEVICES] = {};

static int scratch_setup(int device) {
    if (device < 0 || device >= MAX_DEVICES) {
        set_error("device ordinal out of range for the no-allocation scratch");
        return GFSDC_EINVAL;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return 0;
    void *ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
        set_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        return GFSDC_EHIP;
    }
    g_scratch_ptr[device] = ptr;
    return 0;
}

static void *scratch_ptr(int device) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return (device >= 0 && device < MAX_DEVICES) ? g_scratch_ptr[device] : nullptr;
}

// descriptors of a plain batched synthesis: field f copied to column pair f
__global__ void k_copy_descs(SynthDesc *__restrict__ d, unsigned char *__restrict__ rflag, int nf) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nf) return;
    d[f] = SynthDesc{SK_COPY, f, -1, 1.0, 0.0, 0, 0};
    rflag[f] = 0;
}

static size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }

template <typename Tin, typename Tout>
__global__ void k_convert(const Tin *__restrict__ a, Tout *__restrict__ b, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) b[i] = (Tout)a[i];
}

template <typename Tin, typename Tout> static void convert(const Tin *a, Tout *b, size_t n, hipStream_t s) {
    if (n == 0) return;
    size_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL((k_convert<Tin, Tout>), dim3((unsigned)blocks), dim3(256), 0, s, a, b, n);
}

template <typename T> int Engine<T>::init(const gfsdc_cfg &c) {
    cfg = c;
    if (c.nlev > 0) {
        if (!c.ak || !c.bk) {
            set_error("cfg.ak / cfg.bk required when nlev > 0");
            return GFSDC_EINVAL;
        }
        ak.assign(c.ak, c.ak + c.nlev + 1);
        bk.assign(c.bk, c.bk + c.nlev + 1);
    }
    cfg.ak = ak.empty() ? nullptr : ak.data();
    cfg.bk = bk.empty() ? nullptr : bk.data();
    if (cfg.nmem <= 0) cfg.nmem = 1;
    GFSDC_CHECK(hipSetDevice(c.device));
    hipDeviceProp_t prop;
    GFSDC_CHECK(hipGetDeviceProperties(&prop, c.device));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
        set_error(std::string("gfsdc is built for gfx950; device is ") + prop.gcnArchName);
        return GFSDC_ENODEV;
    }
    GFSDC_CHECK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    int rc = scratch_setup(c.device);
    if (rc) return rc;
    rc = plan_create(plan, c.nlon, c.nlat, c.ntrunc, stream);
    plan.leg_bf16x3 = (c.precision == GFSDC_BF16X3) ? 1 : 0;
    if (rc) return rc;
    if (c.nlev > 0) {
        rc = dyn_init();
        if (rc) return rc;
    }
    return 0;
}

template <typename T> void Engine<T>::destroy() {
    if (stream) (void)hipStreamSynchronize(stream);
    dyn_destroy();
    work_free(work);
    plan_destroy(plan);
    for (int i = 0; i < 4; i++)
        if (stage[i]) (void)hipFree(stage[i]);
    if (d_sdesc) (void)hipFree(d_sdesc);
    if (d_adesc) (void)hipFree(d_adesc);
    if (d_rflag) (void)hipFree(d_rflag);
    if (stream) (void)hipStreamDestroy(stream);
    stream = nullptr;
}

template <typename T> int Engine<T>::ensure_stage(int i, size_t bytes) {
    if (bytes <= stage_sz[i]) return 0;
    GFSDC_CHECK(hipStreamSynchronize(stream));
    if (stage[i]) GFSDC_CHECK(hipFree(stage[i]));
    stage[i] = nullptr;
    stage_sz[i] = 0;
    GFSDC_CHECK(hipMalloc(&stage[i], bytes));
    stage_sz[i] = bytes;
    return 0;
}

template <typename T>
int Engine<T>::upload_descs(const std::vector<SynthDesc> &sd, const std::vector<AnalDesc> &ad,
                            const std::vector<unsigned char> &rf) {
    int need = (int)std::max(sd.size(), std::max(ad.size(), rf.size()));
    if (need > desc_cap) {
        if (d_sdesc) GFSDC_CHECK(hipFree(d_sdesc));
        if (d_adesc) GFSDC_CHECK(hipFree(d_adesc));
        if (d_rflag) GFSDC_CHECK(hipFree(d_rflag));
        GFSDC_CHECK(hipMalloc((void **)&d_sdesc, sizeof(SynthDesc) * need));
        GFSDC_CHECK(hipMalloc((void **)&d_adesc, sizeof(AnalDesc) * need));
        GFSDC_CHECK(hipMalloc((void **)&d_rflag, need));
        desc_cap = need;
    }
    h_sdesc = sd;
    h_adesc = ad;
    h_rflag = rf;
    if (!sd.empty())
        GFSDC_CHECK(hipMemcpyAsync(d_sdesc, h_sdesc.data(), sizeof(SynthDesc) * sd.size(), hipMemcpyHostToDevice,
                                   stream));
    if (!ad.empty())
        GFSDC_CHECK(hipMemcpyAsync(d_adesc, h_adesc.data(), sizeof(AnalDesc) * ad.size(), hipMemcpyHostToDevice,
                                   stream));
    if (!rf.empty())
        GFSDC_CHECK(hipMemcpyAsync(d_rflag, h_rflag.data(), rf.size(), hipMemcpyHostToDevice, stream));
    return 0;
}

template <typename T> int Engine<T>::input(const void *p, size_t n, int dtype, int loc, int slot, const T **out) {
    if (!p) {
        set_error("null buffer");
        return GFSDC_EINVAL;
    }
    if (dtype != GFSDC_DT_F64 && dtype != GFSDC_DT_F32) {
        set_error("bad dtype");
        return GFSDC_EINVAL;
    }
    size_t esz = dtype == GFSDC_DT_F64 ? 8 : 4;
    bool same = (esz == sizeof(T));
    if (loc == GFSDC_DEVICE && same) {
        *out = (const T *)p;
        return 0;
    }
    int rc = ensure_stage(slot, n * (esz + sizeof(T)));
    if (rc) return rc;
    char *raw = (char *)stage[slot];
    T *conv = (T *)(raw + n * esz);
    const void *src = p;
    if (loc == GFSDC_HOST) {
        GFSDC_CHECK(hipMemcpyAsync(raw, p, n * esz, hipMemcpyHostToDevice, stream));
        src = raw;
        if (same) {
            *out = (const T *)raw;
            return 0;
        }
    }
    if (dtype == GFSDC_DT_F64)
        convert((const double *)src, conv, n, stream);
    else
        convert((const float *)src, conv, n, stream);
    GFSDC_CHECK(hipGetLastError());
    *out = conv;
    return 0;
}

template <typename T> int Engine<T>::output(const T *d, void *p, size_t n, int dtype, int loc, int slot) {
    if (!p) {
        set_error("null buffer");
        return GFSDC_EINVAL;
    }
    size_t esz = dtype == GFSDC_DT_F64 ? 8 : 4;
    bool same = (esz == sizeof(T));
    if (loc == GFSDC_DEVICE) {
        if (same)
            GFSDC_CHECK(hipMemcpyAsync(p, d, n * esz, hipMemcpyDeviceToDevice, stream));
        else if (dtype == GFSDC_DT_F64)
            convert(d, (double *)p, n, stream);
        else
            convert(d, (float *)p, n, stream);
        GFSDC_CHECK(hipGetLastError());
        return 0;
    }
    const void *src = d;
    if (!same) {
        int rc = ensure_stage(slot, n * esz);
        if (rc) return rc;
        if (dtype == GFSDC_DT_F64)
            convert(d, (double *)stage[slot], n, stream);
        else
            convert(d, (float *)stage[slot], n, stream);
        src = stage[slot];
    }
    GFSDC_CHECK(hipMemcpyAsync(p, src, n * esz, hipMemcpyDeviceToHost, stream));
    GFSDC_CHECK(hipStreamSynchronize(stream));
    return 0;
}

template <typename T> int Engine<T>::spectogrd(const void *spec, void *grid, int nf, int dtype, int loc) {
    if (nf <= 0) return nf == 0 ? 0 : GFSDC_EINVAL;
    int rc = work_alloc(work, plan, nf);
    if (rc) return rc;
    const T *in;
    if ((rc = input(spec, (size_t)nf * plan.nlm * 2, dtype, loc, 0, &in))) return rc;
    std::vector<SynthDesc> sd(nf);
    for (int f = 0; f < nf; f++) sd[f] = SynthDesc{SK_COPY, f, -1, 1.0, 0.0, 0, 0};
    if ((rc = upload_descs(sd, {}, std::vector<unsigned char>(nf, 0)))) return rc;
    launch_synth_gather(plan, (const cplx<T> *)in, plan.nlm, 1, d_sdesc, nf, work.sx, work.ldc, 1.0, stream);
    launch_leg_synth(plan, work.sx, work.ldc, work.fs, work.ldc, 2 * nf, stream);
    launch_fft_synth(plan, work.fs, work.ldc, work.grid, (long)plan.nlat * plan.nlon, nf, d_rflag, stream);
    GFSDC_CHECK(hipGetLastError());
    return output(work.grid, grid, (size_t)nf * plan.nlat * plan.nlon, dtype, loc, 1);
}

// spectogrd on device buffers of the context's element type with every
// intermediate in the per-device scratch: no allocation, no host copy, no
// synchronisation (work is queued on the context stream).
template <typename T> int Engine<T>::spectogrd_noalloc(const void *spec, void *grid, int nf) {
    if (nf == 0) return 0;
    if (nf < 0 || !spec || !grid) {
        set_error("spectogrd_noalloc: need nfield >= 0 and device buffers");
        return GFSDC_EINVAL;
    }
    char *base = (char *)scratch_ptr(cfg.device);
    if (!base) {
        set_error("spectogrd_noalloc: no scratch address for this device");
        return GFSDC_ESTATE;
    }
    const int ldc = round_up(2 * nf, COL_PAD) + COL_SKEW;
    const size_t o_rf = align256(sizeof(SynthDesc) * nf);
    const size_t o_sx = o_rf + align256((size_t)nf);
    const size_t o_fs = o_sx + align256(sizeof(T) * (size_t)plan.nlme * ldc);
    const size_t need = o_fs + sizeof(T) * (size_t)(plan.ntrunc + 1) * plan.nlat * ldc;
    if (need > NOALLOC_SCRATCH_BYTES) {
        set_error("spectogrd_noalloc: " + std::to_string(nf) + " fields at T" + std::to_string(plan.ntrunc) + " (" +
                  std::to_string(plan.nlon) + "x" + std::to_string(plan.nlat) + ") need " + std::to_string(need >> 20) +
                  " MiB of scratch; the maximum is " + std::to_string(NOALLOC_SCRATCH_BYTES >> 20) +
                  " MiB (use gfsdc_spectogrd, or fewer fields per call)");
        return GFSDC_EINVAL;
    }
    SynthDesc *desc = (SynthDesc *)base;
    unsigned char *rflag = (unsigned char *)(base + o_rf);
    T *sx = (T *)(base + o_sx), *fs = (T *)(base + o_fs);
    hipLaunchKernelGGL(k_copy_descs, dim3((nf + 255) / 256), dim3(256), 0, stream, desc, rflag, nf);
    launch_synth_gather(plan, (const cplx<T> *)spec, plan.nlm, 1, desc, nf, sx, ldc, 1.0, stream);
    launch_leg_synth(plan, sx, ldc, fs, ldc, 2 * nf, stream);
    launch_fft_synth(plan, fs, ldc, (T *)grid, (long)plan.nlat * plan.nlon, nf, rflag, stream);
    GFSDC_CHECK(hipGetLastError());
    return 0;
}

template <typename T> int Engine<T>::grdtospec(const void *grid, void *spec, int nf, int dtype, int loc) {
    if (nf <= 0) return nf == 0 ? 0 : GFSDC_EINVAL;
    int rc = work_alloc(work, plan, nf);
    if (rc) return rc;
    const T *in;
    size_t ng = (size_t)plan.nlat * plan.nlon;
    if ((rc = input(grid, (size_t)nf * ng, dtype, loc, 0, &in))) return rc;
    std::vector<AnalDesc> ad(nf);
    for (int f = 0; f < nf; f++) ad[f] = AnalDesc{AK_COPY, f, -1, 1.0, 0, 0, -1, 0.0};
    if ((rc = upload_descs({}, ad, std::vector<unsigned char>(nf, 0)))) return rc;
    launch_fft_anal(plan, in, (long)ng, nf, d_rflag, work.fa, work.ldc, stream);
    launch_leg_anal(plan, work.fa, work.ldc, work.sx, work.ldc, 2 * nf, stream);
    T *out = work.grid;  // reuse as spectral output buffer (nf * nlm complex <= nf * ng)
    launch_anal_scatter(plan, work.sx, work.ldc, d_adesc, nf, (cplx<T> *)out, plan.nlm, 1, 1.0, stream);
    GFSDC_CHECK(hipGetLastError());
    return output(out, spec, (size_t)nf * plan.nlm * 2, dtype, loc, 1);
}

// Transform microbenchmark with device-resident buffers: per iteration a
// spectogrd + grdtospec round trip of ns scalar fields and a getuv of nv
// vector pairs through the same entry points as callers (loc = device).
template <typename T> __global__ void k_fill_spec(cplx<T> *s, long n, int nlm) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int k = (int)(i % nlm);
    T a = T(1) / (T(1) + T(k % 97));
    s[i] = cplx<T>{a, (k % 7) ? a : T(0)};
}

template <typename T> int Engine<T>::bench_transforms(int ns, int nv, int reps, double *ms) {
    if (ns < 1 || nv < 0 || reps < 1) {
        set_error("bench_transforms: ns >= 1, nv >= 0, reps >= 1");
        return GFSDC_EINVAL;
    }
    const size_t ng = (size_t)plan.nlat * plan.nlon, nlm = plan.nlm;
    const int dt = sizeof(T) == 8 ? GFSDC_DT_F64 : GFSDC_DT_F32;
    T *spec = nullptr, *grid = nullptr, *spec2 = nullptr, *uv = nullptr;
    GFSDC_CHECK(hipMalloc((void **)&spec, sizeof(cplx<T>) * nlm * (ns + 2 * nv)));
    GFSDC_CHECK(hipMalloc((void **)&spec2, sizeof(cplx<T>) * nlm * ns));
    GFSDC_CHECK(hipMalloc((void **)&grid, sizeof(T) * ng * ns));
    GFSDC_CHECK(hipMalloc((void **)&uv, sizeof(T) * ng * 2 * (nv > 0 ? nv : 1)));
    const long nsp = (long)nlm * (ns + 2 * nv);
    hipLaunchKernelGGL(k_fill_spec<T>, dim3((unsigned)((nsp + 255) / 256)), dim3(256), 0, stream, (cplx<T> *)spec,
                       nsp, (int)nlm);
    const T *vrt = spec + 2 * nlm * ns, *div = vrt + 2 * nlm * nv;
    auto iter = [&]() -> int {
        int rc = spectogrd(spec, grid, ns, dt, GFSDC_DEVICE);
        if (!rc) rc = grdtospec(grid, spec2, ns, dt, GFSDC_DEVICE);
        if (!rc && nv > 0)
            rc = sphtor_to_spat(vrt, div, uv, uv + ng * nv, nv, 6.3712e6, 1, 6.3712e6, 1, dt, GFSDC_DEVICE);
        return rc;
    };
    int rc = iter();
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (!rc) {
        GFSDC_CHECK(hipStreamSynchronize(stream));
        GFSDC_CHECK(hipEventCreate(&e0));
        GFSDC_CHECK(hipEventCreate(&e1));
        GFSDC_CHECK(hipEventRecord(e0, stream));
        for (int r = 0; r < reps && !rc; r++) rc = iter();
        GFSDC_CHECK(hipEventRecord(e1, stream));
        GFSDC_CHECK(hipEventSynchronize(e1));
        float t = 0.f;
        GFSDC_CHECK(hipEventElapsedTime(&t, e0, e1));
        if (ms) *ms = t / reps;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(spec);
    (void)hipFree(spec2);
    (void)hipFree(grid);
    (void)hipFree(uv);
    return rc;
}

// plain 16-byte-vector copy, grid-stride
__global__ void k_copy_bytes(float4 *__restrict__ dst, const float4 *__restrict__ src, long n) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

template <typename T> int Engine<T>::bench_copy(long nbytes, int reps, double *ms) {
    if (nbytes < 16 || nbytes % 16 || reps < 1) {
        set_error("bench_copy: nbytes a positive multiple of 16, reps >= 1");
        return GFSDC_EINVAL;
    }
    void *a = nullptr, *b = nullptr;
    GFSDC_CHECK(hipMalloc(&a, nbytes));
    if (hipMalloc(&b, nbytes) != hipSuccess) {
        (void)hipFree(a);
        set_error("bench_copy: device allocation failed");
        return GFSDC_ENOMEM;
    }
    const long n = nbytes / 16;
    GFSDC_CHECK(hipMemsetAsync(a, 0, nbytes, stream));
    hipEvent_t e0, e1;
    GFSDC_CHECK(hipEventCreate(&e0));
    GFSDC_CHECK(hipEventCreate(&e1));
    GFSDC_CHECK(hipEventRecord(e0, stream));
    for (int r = 0; r < reps; r++)
        hipLaunchKernelGGL(k_copy_bytes, dim3(4096), dim3(256), 0, stream, (float4 *)b, (const float4 *)a, n);
    GFSDC_CHECK(hipEventRecord(e1, stream));
    GFSDC_CHECK(hipEventSynchronize(e1));
    float t = 0.f;
    GFSDC_CHECK(hipEventElapsedTime(&t, e0, e1));
    if (ms) *ms = t / reps;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
    return 0;
}

// (S, T) -> (Vt, Vp) of SHTns sphtor_to_spat with S = sscale [invlap] s, T = tscale [invlap] t;
// s may be null (zero spheroidal part, as getgrad).
template <typename T>
int Engine<T>::sphtor_to_spat(const void *s, const void *t, void *u, void *v, int nf, double sscale, int sinv,
                              double tscale, int tinv, int dtype, int loc) {
    if (nf <= 0) return nf == 0 ? 0 : GFSDC_EINVAL;
    int rc = work_alloc(work, plan, 2 * nf);
    if (rc) return rc;
    size_t ns = (size_t)nf * plan.nlm * 2;
    // both inputs in one device buffer [2 nf][nlm]
    if ((rc = ensure_stage(2, 2 * ns * sizeof(T)))) return rc;
    T *both = (T *)stage[2];
    const T *in;
    if (s) {
        if ((rc = input(s, ns, dtype, loc, 0, &in))) return rc;
        GFSDC_CHECK(hipMemcpyAsync(both, in, ns * sizeof(T), hipMemcpyDeviceToDevice, stream));
    } else {
        GFSDC_CHECK(hipMemsetAsync(both, 0, ns * sizeof(T), stream));
    }
    if ((rc = input(t, ns, dtype, loc, 0, &in))) return rc;
    GFSDC_CHECK(hipMemcpyAsync(both + ns, in, ns * sizeof(T), hipMemcpyDeviceToDevice, stream));
    std::vector<SynthDesc> sd(2 * nf);
    for (int f = 0; f < nf; f++) {
        sd[f] = SynthDesc{SK_UHAT, f, nf + f, sscale, tscale, sinv, tinv};
        sd[nf + f] = SynthDesc{SK_VHAT, f, nf + f, sscale, tscale, sinv, tinv};
    }
    if ((rc = upload_descs(sd, {}, std::vector<unsigned char>(2 * nf, 1)))) return rc;
    launch_synth_gather(plan, (const cplx<T> *)both, plan.nlm, 1, d_sdesc, 2 * nf, work.sx, work.ldc, 1.0, stream);
    launch_leg_synth(plan, work.sx, work.ldc, work.fs, work.ldc, 4 * nf, stream);
    size_t ng = (size_t)plan.nlat * plan.nlon;
    launch_fft_synth(plan, work.fs, work.ldc, work.grid, (long)ng, 2 * nf, d_rflag, stream);
    GFSDC_CHECK(hipGetLastError());
    if ((rc = output(work.grid, u, nf * ng, dtype, loc, 1))) return rc;
    return output(work.grid + nf * ng, v, nf * ng, dtype, loc, 1);
}

// (Vt, Vp) -> vorticity / divergence (lap/rsphere S, lap/rsphere T), or SHTns (S, T) when inv = 1.
template <typename T>
int Engine<T>::spat_to_sphtor(const void *u, const void *v, void *s, void *t, int nf, double rsphere, int inv,
                              int dtype, int loc) {
    if (nf <= 0) return nf == 0 ? 0 : GFSDC_EINVAL;
    int rc = work_alloc(work, plan, 2 * nf);
    if (rc) return rc;
    size_t ng = (size_t)plan.nlat * plan.nlon;
    if ((rc = ensure_stage(2, 2 * nf * ng * sizeof(T)))) return rc;
    T *both = (T *)stage[2];
    const T *in;
    if ((rc = input(u, nf * ng, dtype, loc, 0, &in))) return rc;
    GFSDC_CHECK(hipMemcpyAsync(both, in, nf * ng * sizeof(T), hipMemcpyDeviceToDevice, stream));
    if ((rc = input(v, nf * ng, dtype, loc, 0, &in))) return rc;
    GFSDC_CHECK(hipMemcpyAsync(both + nf * ng, in, nf * ng * sizeof(T), hipMemcpyDeviceToDevice, stream));
    std::vector<AnalDesc> ad(2 * nf);
    for (int f = 0; f < nf; f++) {
        ad[f] = AnalDesc{AK_VRT, f, nf + f, 1.0, inv, 0, -1, 0.0};
        ad[nf + f] = AnalDesc{AK_DIV, f, nf + f, 1.0, inv, 0, -1, 0.0};
    }
    if ((rc = upload_descs({}, ad, std::vector<unsigned char>(2 * nf, 1)))) return rc;
    launch_fft_anal(plan, both, (long)ng, 2 * nf, d_rflag, work.fa, work.ldc, stream);
    launch_leg_anal(plan, work.fa, work.ldc, work.sx, work.ldc, 4 * nf, stream);
    T *out = work.grid;
    launch_anal_scatter(plan, work.sx, work.ldc, d_adesc, 2 * nf, (cplx<T> *)out, plan.nlm, 1, rsphere, stream);
    GFSDC_CHECK(hipGetLastError());
    size_t ns = (size_t)nf * plan.nlm * 2;
    if ((rc = output(out, s, ns, dtype, loc, 1))) return rc;
    return output(out + ns, t, ns, dtype, loc, 1);
}

template struct Engine<float>;
template struct Engine<double>;

}  // namespace gfsdc
