// C ABI of libgfsdc: the gfsdc_* handle API and the SHTns symbols that
// src/shtns.f90 binds (shtns.f90:109-112,125-126,160,195,225,261,303).
#include <cmath>
#include <cstring>
#include <mutex>

#include "engine.h"

namespace gfsdc {
static thread_local std::string g_last_error;
void set_error(const std::string &msg) { g_last_error = msg; }
}  // namespace gfsdc

using gfsdc::set_error;

#define DISPATCH(ctx, call)                                         \
    do {                                                            \
        if (!(ctx)) {                                               \
            set_error("null context");                              \
            return GFSDC_EINVAL;                                    \
        }                                                           \
        return (ctx)->ef ? (ctx)->ef->call : (ctx)->ed->call;       \
    } while (0)

extern "C" {

const char *gfsdc_last_error(void) { return gfsdc::g_last_error.c_str(); }

void gfsdc_cfg_defaults(gfsdc_cfg *c) {
    std::memset(c, 0, sizeof(*c));
    c->nmem = 1;
    c->precision = GFSDC_FP64;
    c->physics = GFSDC_PHYS_HS;
    c->massfix = 1;
    c->hdif_fac = 1.0;
    c->hdif_fac2 = 1.0;
    // default-real literals of params.f90 / physcons.f90, rounded as the reference rounds them
    c->taustratdamp = (double)(float)(5. * 86400.);
    c->rerth = (double)6.3712e+6f;
    c->grav = (double)9.80665e+0f;
    c->omega = (double)7.2921e-5f;
    c->rd = (double)2.8705e+2f;
    c->rv = (double)4.6150e+2f;
    c->cp = (double)1.0046e+3f;
    c->cvap = (double)1.8460e+3f;
    c->simple_test = 0;
    c->turbflux = 1;
    c->tpat_amp = 0.0;
    c->tpat_tau = 6.0 * 3600.0;
    c->tpat_lscale = 500.0e3;
    c->tpat_seed = 1;
}

size_t gfsdc_cfg_size(void) { return sizeof(gfsdc_cfg); }

int gfsdc_create(const gfsdc_cfg *cfg, gfsdc_ctx **out) {
    if (!cfg || !out) {
        set_error("null argument");
        return GFSDC_EINVAL;
    }
    *out = nullptr;
    gfsdc_ctx *c = new gfsdc_ctx();
    c->precision = cfg->precision;
    int rc;
    if (cfg->precision == GFSDC_FP32 || cfg->precision == GFSDC_BF16X3) {
        c->ef = new gfsdc::Engine<float>();
        rc = c->ef->init(*cfg);
    } else if (cfg->precision == GFSDC_FP64) {
        c->ed = new gfsdc::Engine<double>();
        rc = c->ed->init(*cfg);
    } else {
        set_error("precision must be GFSDC_FP64, GFSDC_FP32 or GFSDC_BF16X3");
        rc = GFSDC_EINVAL;
    }
    if (rc) {
        gfsdc_destroy(c);
        return rc;
    }
    *out = c;
    return 0;
}

int gfsdc_destroy(gfsdc_ctx *c) {
    if (!c) return 0;
    if (c->ef) {
        c->ef->destroy();
        delete c->ef;
    }
    if (c->ed) {
        c->ed->destroy();
        delete c->ed;
    }
    delete c;
    return 0;
}

int gfsdc_nlm(const gfsdc_ctx *c) {
    if (!c) return GFSDC_EINVAL;
    return c->ef ? c->ef->plan.nlm : c->ed->plan.nlm;
}

int gfsdc_griddata(const gfsdc_ctx *c, double *lats, double *gauwts, double *areawts, int *degree, int *order,
                   double *lap, double *invlap) {
    if (!c) {
        set_error("null context");
        return GFSDC_EINVAL;
    }
    const std::vector<double> &mu = c->ef ? c->ef->plan.mu : c->ed->plan.mu;
    const std::vector<double> &w = c->ef ? c->ef->plan.w : c->ed->plan.w;
    int nlon = c->ef ? c->ef->plan.nlon : c->ed->plan.nlon;
    int nlat = (int)mu.size();
    int T = c->ef ? c->ef->plan.ntrunc : c->ed->plan.ntrunc;
    double tot = 0.0;
    for (int j = 0; j < nlat; j++) tot += w[j] * nlon;
    for (int j = 0; j < nlat; j++) {
        if (lats) lats[j] = std::asin(mu[j]);
        if (gauwts) gauwts[j] = w[j];
        if (areawts)
            for (int i = 0; i < nlon; i++) areawts[(size_t)j * nlon + i] = w[j] / tot;
    }
    int k = 0;
    for (int m = 0; m <= T; m++)
        for (int n = m; n <= T; n++, k++) {
            if (degree) degree[k] = n;
            if (order) order[k] = m;
            if (lap) lap[k] = -(double)n * (n + 1.0);
            if (invlap) invlap[k] = k == 0 ? 0.0 : -1.0 / ((double)n * (n + 1.0));
        }
    return 0;
}

int gfsdc_spectogrd(gfsdc_ctx *c, const void *spec, void *grid, int nf, int dtype, int loc) {
    DISPATCH(c, spectogrd(spec, grid, nf, dtype, loc));
}
int gfsdc_grdtospec(gfsdc_ctx *c, const void *grid, void *spec, int nf, int dtype, int loc) {
    DISPATCH(c, grdtospec(grid, spec, nf, dtype, loc));
}
int gfsdc_spectogrd_noalloc(gfsdc_ctx *c, const void *spec, void *grid, int nf) {
    DISPATCH(c, spectogrd_noalloc(spec, grid, nf));
}
// shtns.f90:259-261: S = invlap*rsphere*vrt, T = invlap*rsphere*div; u = Vt, v = Vp
int gfsdc_getuv(gfsdc_ctx *c, const void *vrt, const void *div, void *u, void *v, int nf, double rsphere, int dtype,
                int loc) {
    DISPATCH(c, sphtor_to_spat(vrt, div, u, v, nf, rsphere, 1, rsphere, 1, dtype, loc));
}
// shtns.f90:303,310: spat_to_sphtor then * lap / rsphere
int gfsdc_getvrtdivspec(gfsdc_ctx *c, const void *u, const void *v, void *vrt, void *div, int nf, double rsphere,
                        int dtype, int loc) {
    DISPATCH(c, spat_to_sphtor(u, v, vrt, div, nf, rsphere, 0, dtype, loc));
}
// shtns.f90:339-351: sphtor_to_spat(0, f) / rsphere
int gfsdc_getgrad(gfsdc_ctx *c, const void *spec, void *gx, void *gy, int nf, double rsphere, int dtype, int loc) {
    DISPATCH(c, sphtor_to_spat(nullptr, spec, gx, gy, nf, 0.0, 0, 1.0 / rsphere, 0, dtype, loc));
}

int gfsdc_set_state(gfsdc_ctx *c, const void *vrt, const void *div, const void *tv, const void *trac,
                    const void *lnps, int dtype, int loc) {
    DISPATCH(c, set_state(vrt, div, tv, trac, lnps, dtype, loc));
}
int gfsdc_get_state(gfsdc_ctx *c, void *vrt, void *div, void *tv, void *trac, void *lnps, int dtype, int loc) {
    DISPATCH(c, get_state(vrt, div, tv, trac, lnps, dtype, loc));
}
int gfsdc_set_topo(gfsdc_ctx *c, const void *topo, int dtype, int loc) { DISPATCH(c, set_topo(topo, dtype, loc)); }
int gfsdc_set_pdryini(gfsdc_ctx *c, const double *p) { DISPATCH(c, set_pdryini(p)); }
int gfsdc_get_pdryini(gfsdc_ctx *c, double *p) { DISPATCH(c, get_pdryini(p)); }
int gfsdc_getdyntend(gfsdc_ctx *c, int rkstage, int inverse_only, void *dvrt, void *ddiv, void *dtv, void *dtrac,
                     void *dlnps, int dtype, int loc) {
    DISPATCH(c, getdyntend(rkstage, inverse_only, dvrt, ddiv, dtv, dtrac, dlnps, dtype, loc));
}
int gfsdc_advance(gfsdc_ctx *c, int nsteps, int use_graph) { DISPATCH(c, advance(nsteps, use_graph)); }
int gfsdc_hs_phytend(gfsdc_ctx *c, void *dvrt, void *ddiv, void *dtv, int dtype, int loc) {
    DISPATCH(c, hs_phytend(dvrt, ddiv, dtv, dtype, loc));
}
int gfsdc_hs_drag_levels(gfsdc_ctx *c, int *n) { DISPATCH(c, hs_drag_levels(n)); }
int gfsdc_simple_phytend(gfsdc_ctx *c, void *dvrt, void *ddiv, void *dtv, void *dtrac, int dtype, int loc) {
    DISPATCH(c, simple_phytend(dvrt, ddiv, dtv, dtrac, dtype, loc));
}
int gfsdc_get_pattern(gfsdc_ctx *c, void *spec, int dtype, int loc) { DISPATCH(c, get_pattern(spec, dtype, loc)); }
int gfsdc_autotune(gfsdc_ctx *c, double *ms, int *best) { DISPATCH(c, autotune(ms, best)); }
int gfsdc_bench_transforms(gfsdc_ctx *c, int nscalar, int nvector, int reps, double *ms) {
    DISPATCH(c, bench_transforms(nscalar, nvector, reps, ms));
}
int gfsdc_bench_copy(gfsdc_ctx *c, long nbytes, int reps, double *ms) { DISPATCH(c, bench_copy(nbytes, reps, ms)); }
int gfsdc_mlp_init(gfsdc_ctx *c, int hidden, int nlayers, int maxbatch, long capacity, unsigned long long seed) {
    DISPATCH(c, mlp_init(hidden, nlayers, maxbatch, capacity, seed));
}
int gfsdc_mlp_collect(gfsdc_ctx *c, int nsteps, int stride, long *ncollected) {
    DISPATCH(c, mlp_collect(nsteps, stride, ncollected));
}
int gfsdc_mlp_normalize(gfsdc_ctx *c) { DISPATCH(c, mlp_normalize()); }
int gfsdc_mlp_train(gfsdc_ctx *c, int nsteps, int batch, double lr, double wd, long ntrain, double *losses) {
    DISPATCH(c, mlp_train(nsteps, batch, lr, wd, ntrain, losses));
}
int gfsdc_mlp_eval(gfsdc_ctx *c, long first, long count, double *err, double *base) {
    DISPATCH(c, mlp_eval(first, count, err, base));
}
int gfsdc_mlp_enable(gfsdc_ctx *c, int on, int judge_stride, double threshold) {
    DISPATCH(c, mlp_enable(on, judge_stride, threshold));
}
int gfsdc_mlp_judge_stats(gfsdc_ctx *c, double *out) { DISPATCH(c, mlp_judge_stats(out)); }
long gfsdc_mlp_nparams(gfsdc_ctx *c) {
    if (!c) return 0;
    return c->ef ? c->ef->mlp_nparams() : (c->ed ? c->ed->mlp_nparams() : 0);
}
int gfsdc_mlp_params(gfsdc_ctx *c, float *get, const float *set) { DISPATCH(c, mlp_params(get, set)); }
int gfsdc_mlp_norm(gfsdc_ctx *c, float *get, const float *set) { DISPATCH(c, mlp_norm(get, set)); }
int gfsdc_mlp_forward(gfsdc_ctx *c, const float *x, long n, float *y) { DISPATCH(c, mlp_forward_host(x, n, y)); }
int gfsdc_mlp_grad(gfsdc_ctx *c, const float *x, const float *t, int n, float *grad, double *loss) {
    DISPATCH(c, mlp_grad_host(x, t, n, grad, loss));
}
int gfsdc_mlp_adamw(gfsdc_ctx *c, const float *grad, double lr, double wd) {
    DISPATCH(c, mlp_adamw_host(grad, lr, wd));
}
int gfsdc_get_grid(gfsdc_ctx *c, int which, void *out, int dtype, int loc) {
    DISPATCH(c, get_grid(which, out, dtype, loc));
}
int gfsdc_get_semimp(gfsdc_ctx *c, double *a, double *b, double *sv, double *tor, double *d) {
    DISPATCH(c, get_semimp(a, b, sv, tor, d));
}
int gfsdc_get_damp(gfsdc_ctx *c, double *disspec, double *diff_prof, double *dmp_prof) {
    DISPATCH(c, get_damp(disspec, diff_prof, dmp_prof));
}
int gfsdc_step_diag(gfsdc_ctx *c, double *out) { DISPATCH(c, step_diag(out)); }
int gfsdc_init_rest_state(gfsdc_ctx *c, double t0, double ps0, double vort_rms, int nmin, int nmax,
                          unsigned long long seed) {
    DISPATCH(c, init_rest_state(t0, ps0, vort_rms, nmin, nmax, seed));
}
int gfsdc_profile_step(gfsdc_ctx *c, char names[][48], double *ms, int nmax) {
    DISPATCH(c, profile_step(names, ms, nmax));
}
int gfsdc_set_time(gfsdc_ctx *c, double t) { DISPATCH(c, set_time(t)); }
int gfsdc_get_time(gfsdc_ctx *c, double *t) { DISPATCH(c, get_time(t)); }
int gfsdc_iau_set(gfsdc_ctx *c, int nfiles, const double *fhrs, double delthrs, const void *vrt, const void *div,
                  const void *tv, const void *trac, const void *lnps, int dtype, int loc) {
    DISPATCH(c, iau_set(nfiles, fhrs, delthrs, vrt, div, tv, trac, lnps, dtype, loc));
}
int gfsdc_dfi(gfsdc_ctx *c, int ntdfi, int use_graph) { DISPATCH(c, dfi(ntdfi, use_graph)); }
void gfsdc_stoch_defaults(gfsdc_stoch_cfg *s) {
    std::memset(s, 0, sizeof(*s));
    s->addnoise_vrtonly = 1;
}
int gfsdc_stoch_init(gfsdc_ctx *c, const gfsdc_stoch_cfg *s) { DISPATCH(c, stoch_init(s)); }
int gfsdc_get_stoch(gfsdc_ctx *c, int which, void *out, int dtype, int loc) {
    DISPATCH(c, get_stoch(which, out, dtype, loc));
}
int gfsdc_checkpoint_size(gfsdc_ctx *c, size_t *nbytes) {
    if (!c || !nbytes) {
        set_error("null argument");
        return GFSDC_EINVAL;
    }
    *nbytes = c->ef ? c->ef->checkpoint_size() : c->ed->checkpoint_size();
    if (!*nbytes) {
        set_error("context has no dycore");
        return GFSDC_ESTATE;
    }
    return 0;
}
int gfsdc_checkpoint_save(gfsdc_ctx *c, void *buf, size_t nbytes, int loc) {
    DISPATCH(c, checkpoint_save(buf, nbytes, loc));
}
int gfsdc_checkpoint_load(gfsdc_ctx *c, const void *buf, size_t nbytes, int loc) {
    DISPATCH(c, checkpoint_load(buf, nbytes, loc));
}
int gfsdc_sync(gfsdc_ctx *c) {
    if (!c) {
        set_error("null context");
        return GFSDC_EINVAL;
    }
    hipError_t e = hipStreamSynchronize(c->ef ? c->ef->stream : c->ed->stream);
    if (e != hipSuccess) {
        set_error(hipGetErrorString(e));
        return GFSDC_EHIP;
    }
    return 0;
}
void *gfsdc_stream(gfsdc_ctx *c) { return c ? (void *)(c->ef ? c->ef->stream : c->ed->stream) : nullptr; }

// ---------------------------------------------------------------------------
// SHTns symbols. SHTns keeps one global configuration set by set_size /
// precompute; so does this layer (an FP64 context). Per-call host arrays, one
// field per call, synchronous. Calls are serialised by a mutex because
// shtns.f90 is called from inside OpenMP loops over levels.
static std::mutex g_shtns_mu;
static gfsdc_ctx *g_shtns_ctx = nullptr;
static int g_lmax = -1, g_mmax = -1, g_mres = 1, g_norm = 0, g_nthreads = 1;
static int g_nlat = 0, g_nphi = 0;

// Errors: the SHTns symbols return nothing, so an error cannot travel back
// through them. In the default mode (0) the layer prints the error and ends
// the process with exit status 1 -- what SHTns' own error path and the
// reference wrapper's `stop` do. In mode 1 it records the error and returns;
// gfsdc_shtns_status() then reports (and clears) it, and the outputs of the
// failed call are unchanged.
static int g_shtns_mode = 0;
static int g_shtns_status = 0;

static void shtns_fail(const char *what) {
    std::string msg = std::string(what) + ": " + gfsdc::g_last_error;
    g_shtns_status = GFSDC_ESTATE;
    if (g_shtns_mode == 0) {
        fprintf(stderr, "gfsdc shtns layer: %s\n", msg.c_str());
        fflush(stderr);
        exit(1);
    }
    gfsdc::g_last_error = msg;
}

static bool shtns_ready(const char *what) {
    if (g_shtns_ctx) return true;
    gfsdc::g_last_error = "no SHTns configuration (shtns_precompute_ not called)";
    shtns_fail(what);
    return false;
}

void gfsdc_shtns_error_mode(int mode) {
    std::lock_guard<std::mutex> lk(g_shtns_mu);
    g_shtns_mode = mode ? 1 : 0;
}

int gfsdc_shtns_status(void) {
    std::lock_guard<std::mutex> lk(g_shtns_mu);
    int s = g_shtns_status;
    g_shtns_status = 0;
    return s;
}

void shtns_use_threads_(int *nthreads) { g_nthreads = nthreads ? *nthreads : 1; }

void shtns_set_size_(int *lmax, int *mmax, int *mres, int *norm) {
    std::lock_guard<std::mutex> lk(g_shtns_mu);
    g_lmax = *lmax;
    g_mmax = *mmax;
    g_mres = *mres;
    g_norm = *norm;
    // SHT_ORTHONORMAL (0) + SHT_NO_CS_PHASE (1024), triangular truncation, mres 1
    if (g_mmax != g_lmax || g_mres != 1 || (g_norm & 0xFF) != 0 || !(g_norm & 1024)) {
        gfsdc::g_last_error = "only lmax == mmax, mres 1, orthonormal without CS phase is supported";
        shtns_fail("shtns_set_size");
    }
}

void shtns_precompute_(int *type, int *layout, double *eps, int *nlat, int *nphi) {
    (void)eps;
    std::lock_guard<std::mutex> lk(g_shtns_mu);
    // SHT_GAUSS (0) or SHT_GAUSS_FLY (6); SHT_PHI_CONTIGUOUS (512)
    if ((*type != 0 && *type != 6) || (*layout & 512) == 0) {
        gfsdc::g_last_error = "only Gauss grids (SHT_GAUSS, SHT_GAUSS_FLY) in the phi-contiguous layout";
        shtns_fail("shtns_precompute");
        return;
    }
    if (g_lmax < 0) {
        gfsdc::g_last_error = "shtns_set_size_ must come first";
        shtns_fail("shtns_precompute");
        return;
    }
    if (g_shtns_ctx) gfsdc_destroy(g_shtns_ctx);
    g_shtns_ctx = nullptr;
    gfsdc_cfg c;
    gfsdc_cfg_defaults(&c);
    c.nlon = *nphi;
    c.nlat = *nlat;
    c.ntrunc = g_lmax;
    c.nlev = 0;
    c.precision = GFSDC_FP64;
    if (gfsdc_create(&c, &g_shtns_ctx) != 0) {
        g_shtns_ctx = nullptr;
        shtns_fail("shtns_precompute");
        return;
    }
    g_nlat = *nlat;
    g_nphi = *nphi;
}

void shtns_calc_nlm_(int *nlm, int *lmax, int *mmax, int *mres) {
    int L = *lmax, M = *mmax / *mres;
    *nlm = (M + 1) * (L + 1) - (*mres) * M * (M + 1) / 2;
}

void shtns_cos_array_(double *costh) {
    std::lock_guard<std::mutex> lk(g_shtns_mu);
    if (!shtns_ready("shtns_cos_array")) return;
    const std::vector<double> &mu = g_shtns_ctx->ed->plan.mu;
    for (int j = 0; j < g_nlat; j++) costh[j] = mu[j];
}

void shtns_gauss_wts_(double *wts) {
    std::lock_guard<std::mutex> lk(g_shtns_mu);
    if (!shtns_ready("shtns_gauss_wts")) return;
    const std::vector<double> &w = g_shtns_ctx->ed->plan.w;
    for (int j = 0; j < g_nlat / 2; j++) wts[j] = w[j];
}

void shtns_spat_to_sh_(double *vr, double *qlm) {
    std::lock_guard<std::mutex> lk(g_shtns_mu);
    if (!shtns_ready("shtns_spat_to_sh")) return;
    if (gfsdc_grdtospec(g_shtns_ctx, vr, qlm, 1, GFSDC_DT_F64, GFSDC_HOST)) shtns_fail("shtns_spat_to_sh");
}

void shtns_sh_to_spat_(double *qlm, double *vr) {
    std::lock_guard<std::mutex> lk(g_shtns_mu);
    if (!shtns_ready("shtns_sh_to_spat")) return;
    if (gfsdc_spectogrd(g_shtns_ctx, qlm, vr, 1, GFSDC_DT_F64, GFSDC_HOST)) shtns_fail("shtns_sh_to_spat");
}

void shtns_sphtor_to_spat_(double *slm, double *tlm, double *vt, double *vp) {
    std::lock_guard<std::mutex> lk(g_shtns_mu);
    if (!shtns_ready("shtns_sphtor_to_spat")) return;
    if (g_shtns_ctx->ed->sphtor_to_spat(slm, tlm, vt, vp, 1, 1.0, 0, 1.0, 0, GFSDC_DT_F64, GFSDC_HOST))
        shtns_fail("shtns_sphtor_to_spat");
}

void shtns_spat_to_sphtor_(double *vt, double *vp, double *slm, double *tlm) {
    std::lock_guard<std::mutex> lk(g_shtns_mu);
    if (!shtns_ready("shtns_spat_to_sphtor")) return;
    if (g_shtns_ctx->ed->spat_to_sphtor(vt, vp, slm, tlm, 1, 1.0, 1, GFSDC_DT_F64, GFSDC_HOST))
        shtns_fail("shtns_spat_to_sphtor");
}

void shtns_reset_(void) {
    std::lock_guard<std::mutex> lk(g_shtns_mu);
    if (g_shtns_ctx) gfsdc_destroy(g_shtns_ctx);
    g_shtns_ctx = nullptr;
}

}  // extern "C"
