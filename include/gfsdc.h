/*
 * gfsdc -- MI355X (gfx950) spectral transform-and-tendency step of the GFS
 * hydrostatic spectral dynamical core.
 *
 * Plain C ABI (extern "C", pointers and sizes, no torch or HIP types). Two
 * tiers:
 *
 *  1. SHTns-compatible symbols. src/shtns.f90 binds eleven Fortran-callable
 *     C symbols of the SHTns library (all arguments by reference). Linking
 *     the unchanged shtns.f90 against this library instead of
 *     "-lshtns -lfftw3" routes every per-field transform of the reference
 *     to the GPU. Arrays are host FP64, one field per call, synchronous.
 *
 *  2. gfsdc_* handle API. Device-resident, batched, precision-selectable
 *     (FP64 / FP32 MFMA Legendre), used by a Fortran host's getdyntend /
 *     advance replacement so that the state never leaves HBM inside a step.
 *
 * Conventions (as src/shtns.f90, src/spectral_data.f90, src/grid_data.f90):
 *   spectral: complex, (ndimspec, nfield) column-major, m-major index
 *             k(m,n) = m(2T+3-m)/2 + (n-m), orthonormal, no Condon-Shortley
 *             phase; real field f = sum q_n0 Y_n0 + 2 Re sum_{m>0} q_nm Y_nm
 *   grid:     real, (nlon, nlat, nfield) column-major, lon fastest,
 *             latitudes north -> south on Gauss nodes, lon_i = 2 pi i / nlon
 *   levels:   model levels bottom -> top; ak/bk interfaces top -> bottom
 *
 * Errors: every gfsdc_* call returns 0 on success and a negative GFSDC_E*
 * code otherwise; gfsdc_last_error() returns the text of the calling
 * thread's last error. Calls on one context are not thread-safe; use one
 * context per host thread or serialise.
 */
#ifndef GFSDC_H
#define GFSDC_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ */
/* Tier 1: SHTns symbols bound by src/shtns.f90                        */
/* ------------------------------------------------------------------ */
/* replaces SHTns shtns_use_threads (shtns.f90:109); the count is recorded, GPU work ignores it */
void shtns_use_threads_(int *nthreads);
/* replaces SHTns shtns_set_size (shtns.f90:110): lmax, mmax, mres, norm flags */
void shtns_set_size_(int *lmax, int *mmax, int *mres, int *norm);
/* replaces SHTns shtns_precompute (shtns.f90:111): grid type, layout, polar threshold, nlat, nphi */
void shtns_precompute_(int *type, int *layout, double *eps, int *nlat, int *nphi);
/* replaces SHTns shtns_calc_nlm (shtns.f90:112) */
void shtns_calc_nlm_(int *nlm, int *lmax, int *mmax, int *mres);
/* replaces SHTns shtns_cos_array (shtns.f90:125): cos(colatitude), nlat values, N->S */
void shtns_cos_array_(double *costh);
/* replaces SHTns shtns_gauss_wts (shtns.f90:126): nlat/2 Gauss weights (full set sums to 2) */
void shtns_gauss_wts_(double *wts);
/* replaces SHTns shtns_spat_to_sh (shtns.f90:195,199) */
void shtns_spat_to_sh_(double *vr, double *qlm /* complex */);
/* replaces SHTns shtns_sh_to_spat (shtns.f90:225,229) */
void shtns_sh_to_spat_(double *qlm /* complex */, double *vr);
/* replaces SHTns shtns_sphtor_to_spat (shtns.f90:261,270,341,348) */
void shtns_sphtor_to_spat_(double *slm, double *tlm, double *vt, double *vp);
/* replaces SHTns shtns_spat_to_sphtor (shtns.f90:303,308) */
void shtns_spat_to_sphtor_(double *vt, double *vp, double *slm, double *tlm);
/* replaces SHTns shtns_reset (shtns.f90:160) */
void shtns_reset_(void);
/* Errors of the symbols above (no counterpart in SHTns, whose symbols return
 * nothing): in mode 0 (default) an error is printed and the process exits
 * with status 1, as SHTns' own error path and the wrapper's `stop` do; in
 * mode 1 it is recorded and the call returns with its outputs unchanged, and
 * gfsdc_shtns_status() returns that error's code (0 if none) and clears it.
 * gfsdc_last_error() holds its text. For callers that must not lose the process. */
void gfsdc_shtns_error_mode(int mode);
int gfsdc_shtns_status(void);

/* ------------------------------------------------------------------ */
/* Tier 2: gfsdc handle API                                            */
/* ------------------------------------------------------------------ */
enum {
    GFSDC_OK = 0,
    GFSDC_EINVAL = -1,   /* bad argument / shape */
    GFSDC_EHIP = -2,     /* HIP runtime error */
    GFSDC_ENOMEM = -3,   /* device allocation failed */
    GFSDC_ESTATE = -4,   /* call out of sequence */
    GFSDC_ENODEV = -5    /* no gfx950 device */
};

/* compute precision of the context: FP64 throughout; FP32 (FP64 tables and
 * global sums); BF16X3 = FP32 with the Legendre GEMMs as three bf16 matrix-core
 * products per FP32 product (split operands hi + lo, lo.lo dropped) */
enum { GFSDC_FP64 = 0, GFSDC_FP32 = 1, GFSDC_BF16X3 = 2 };
enum { GFSDC_HOST = 0, GFSDC_DEVICE = 1 };           /* location of a caller buffer */
enum { GFSDC_DT_F64 = 0, GFSDC_DT_F32 = 1 };         /* element type of a caller buffer */
enum { GFSDC_PHYS_NONE = 0, GFSDC_PHYS_HS = 1, GFSDC_PHYS_SIMPLE = 2 }; /* physics plug-in (Makefile:50-72) */

typedef struct gfsdc_cfg {
    int nlon, nlat, ntrunc, nlev, ntrac, nmem;  /* nmem: ensemble members folded into one launch */
    int precision;                              /* GFSDC_FP64 or GFSDC_FP32 */
    double dt;                                  /* seconds (params.f90 deltim) */
    const double *ak, *bk;                      /* nlev+1 interfaces, top -> bottom, Pa / unitless */
    int explicit_;                              /* params.f90 explicit (0 = IMEX semi-implicit) */
    int physics;                                /* GFSDC_PHYS_* */
    int massfix;                                /* params.f90 massfix */
    int ndiss;                                  /* 0 -> 8 (dyn_init.f90:461) */
    double efold;                               /* <= 0 -> GFS default by truncation */
    double hdif_fac, hdif_fac2, fshk, taustratdamp;
    /* physical constants (physcons.f90); gfsdc_cfg_defaults fills them */
    double rerth, grav, omega, rd, rv, cp, cvap;
    int device;                                 /* HIP device ordinal */
    /* simple moist physics (GFSDC_PHYS_SIMPLE, needs ntrac >= 1, tracer 1 = q):
     * simple_test selects the sea-surface temperature -- 0 constant 302.15 K
     * (reference test 0, dcmip 5x), 1 moist baroclinic-wave formula (test 1,
     * dcmip 4x), 2 302.15 K - 30 K sin^2(lat) (this library's own);
     * turbflux: surface fluxes + boundary-layer diffusion (dcmip 43 and 5x) */
    int simple_test, turbflux;
    /* AR(1) spectral temperature pattern per member (patterngenerator.f90,
     * advanced at the start of each step as run.f90:272-292) applied as the
     * temperature forcing pattern/tau at the end of the step, tapered to 0
     * above sigma 0.05 (stoch_data.f90). tpat_amp (K) = 0 disables it;
     * tpat_tau (s), tpat_lscale (m); member i uses seed tpat_seed + i. */
    double tpat_amp, tpat_tau, tpat_lscale;
    int tpat_seed;
} gfsdc_cfg;

typedef struct gfsdc_ctx gfsdc_ctx;

/* fill defaults of params.f90 / physcons.f90 (sizes left 0) */
void gfsdc_cfg_defaults(gfsdc_cfg *cfg);
/* sizeof(gfsdc_cfg) as this library was built (no counterpart in the
 * reference): a binding in another language (the Fortran bind(c) type, a
 * ctypes Structure) checks its own layout against it before calling */
size_t gfsdc_cfg_size(void);
/* replaces init_dyn's shtns_init (dyn_init.f90:43), setdampspec (:86) and
 * init_semimpdata (:88): builds Gauss grid, Legendre tables, damping and
 * semi-implicit operators on the device */
int gfsdc_create(const gfsdc_cfg *cfg, gfsdc_ctx **out);
int gfsdc_destroy(gfsdc_ctx *ctx);
const char *gfsdc_last_error(void);
/* number of spectral coefficients (ndimspec) and the shtns.f90 module data */
int gfsdc_nlm(const gfsdc_ctx *ctx);
int gfsdc_griddata(const gfsdc_ctx *ctx, double *lats, double *gauwts, double *areawts,
                   int *degree, int *order, double *lap, double *invlap);

/* batched per-field transforms (shtns.f90 API over nfield fields; one call
 * replaces an OpenMP loop over levels). spec: (ndimspec, nfield) complex of
 * dtype; grid: (nlon, nlat, nfield) of dtype; loc GFSDC_HOST / GFSDC_DEVICE. */
/* replaces shtns.f90:203 spectogrd */
int gfsdc_spectogrd(gfsdc_ctx *ctx, const void *spec, void *grid, int nfield, int dtype, int loc);
/* No counterpart in the reference: spectogrd for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc). spec and grid are device buffers of the context's
 * element type (double for GFSDC_FP64, float otherwise), layouts as
 * gfsdc_spectogrd. Every intermediate lives in one per-GPU static scratch of
 * 128 MiB (enough for 64 fields at T254 in FP32, or at T126 in FP64); a call
 * that needs more returns GFSDC_EINVAL with the sizes in gfsdc_last_error().
 * The call allocates nothing, copies nothing to or from the host and does not
 * synchronise: the work is queued on the context stream (gfsdc_sync waits).
 * All calls on one GPU share its scratch, so calls that overlap in time on one
 * GPU (two contexts, two streams) race; serialise them. */
int gfsdc_spectogrd_noalloc(gfsdc_ctx *ctx, const void *spec, void *grid, int nfield);
/* replaces shtns.f90:171 grdtospec */
int gfsdc_grdtospec(gfsdc_ctx *ctx, const void *grid, void *spec, int nfield, int dtype, int loc);
/* replaces shtns.f90:233 getuv */
int gfsdc_getuv(gfsdc_ctx *ctx, const void *vrt, const void *div, void *u, void *v, int nfield,
                double rsphere, int dtype, int loc);
/* replaces shtns.f90:275 getvrtdivspec */
int gfsdc_getvrtdivspec(gfsdc_ctx *ctx, const void *u, const void *v, void *vrt, void *div,
                        int nfield, double rsphere, int dtype, int loc);
/* replaces shtns.f90:313 getgrad */
int gfsdc_getgrad(gfsdc_ctx *ctx, const void *spec, void *gx, void *gy, int nfield,
                  double rsphere, int dtype, int loc);

/* model state, member-major: vrt/div/tv (ndimspec, nlev, nmem), trac
 * (ndimspec, nlev, ntrac, nmem), lnps (ndimspec, nmem) complex of dtype.
 * Any pointer may be NULL to leave that field unchanged / unread. */
int gfsdc_set_state(gfsdc_ctx *ctx, const void *vrt, const void *div, const void *tv,
                    const void *trac, const void *lnps, int dtype, int loc);
int gfsdc_get_state(gfsdc_ctx *ctx, void *vrt, void *div, void *tv, void *trac, void *lnps,
                    int dtype, int loc);
/* spectral orography (ndimspec, nmem) in metres; grad(g*topo) as dyn_init.f90:84 */
int gfsdc_set_topo(gfsdc_ctx *ctx, const void *topo, int dtype, int loc);
/* params.f90 pdryini (Pa) per member; 0 means "take from the first stage-1 call" */
int gfsdc_set_pdryini(gfsdc_ctx *ctx, const double *pdryini);
int gfsdc_get_pdryini(gfsdc_ctx *ctx, double *pdryini);

/* replaces dyn_run.f90:33 getdyntend on the context state; tendencies in the
 * state layout (dtrac may be NULL when ntrac == 0, all may be NULL when
 * inverse_only, which then only refreshes the grid fields) */
int gfsdc_getdyntend(gfsdc_ctx *ctx, int rkstage, int inverse_only, void *dvrt, void *ddiv,
                     void *dtv, void *dtrac, void *dlnps, int dtype, int loc);
/* replaces run.f90:194 advance, nsteps times (device resident; captured in a
 * HIP graph after the first step when use_graph != 0). With Held-Suarez
 * physics every step but the last runs a lean physics pass that synthesises
 * only the grids the forcing reads (temperature, ln ps, and vorticity and
 * divergence on the levels where the drag can act: gfsdc_hs_drag_levels);
 * the state is the same bit for bit, and the last step leaves every grid field
 * for gfsdc_get_grid and gfsdc_step_diag (run.f90:157-162). */
int gfsdc_advance(gfsdc_ctx *ctx, int nsteps, int use_graph);
/* Held-Suarez tendencies on the current grid fields (phy_run_hs.f90:22) */
int gfsdc_hs_phytend(gfsdc_ctx *ctx, void *dvrt, void *ddiv, void *dtv, int dtype, int loc);
/* levels, counted from the lowest, on which the last lean Held-Suarez physics
 * pass of gfsdc_advance formed and transformed the drag (0 before the first):
 * on every level above them p / ps stayed below phy_run_hs.f90's sigbot in
 * every column, so their drag and its vorticity and divergence tendencies are
 * zero. A diagnostic with no counterpart in the reference. */
int gfsdc_hs_drag_levels(gfsdc_ctx *ctx, int *nlev);
/* replaces phy_run_simple.f90:28 getphytend (simple moist physics on the current
 * grid fields, i.e. after gfsdc_getdyntend(ctx, 3, 1, ...)): tendencies of
 * vrt, div, tv (ndimspec, nlev, nmem) and trac (ndimspec, nlev, ntrac, nmem);
 * also refreshes the precip and pwat grids. The state is not changed. */
int gfsdc_simple_phytend(gfsdc_ctx *ctx, void *dvrt, void *ddiv, void *dtv, void *dtrac, int dtype,
                         int loc);
/* tile autotuner (no counterpart in the reference): times the step's Fourier
 * transforms with 2, 4 and 8 fields per FFT workgroup on this context's sizes
 * and keeps the fastest, for the synthesis and the analysis separately; ms (3,
 * may be NULL) receives each candidate's synthesis + analysis time, best the
 * synthesis choice. The model state is unchanged. */
int gfsdc_autotune(gfsdc_ctx *ctx, double *ms, int *best);
/* transform microbenchmark (no counterpart in the reference): milliseconds per
 * iteration of a spectogrd + grdtospec round trip of nscalar fields plus a
 * getuv of nvector pairs on device-resident buffers, averaged over reps */
int gfsdc_bench_transforms(gfsdc_ctx *ctx, int nscalar, int nvector, int reps, double *ms);
/* device-to-device copy of nbytes (a multiple of 16) between two fresh buffers,
 * reps times (no counterpart in the reference): the kernel of known size against
 * which profiler memory counters (FETCH_SIZE, WRITE_SIZE) are scaled to bytes;
 * milliseconds per copy */
int gfsdc_bench_copy(gfsdc_ctx *ctx, long nbytes, int reps, double *ms);
/* current AR(1) pattern, (ndimspec, nmem) complex of dtype (error if tpat_amp == 0) */
int gfsdc_get_pattern(gfsdc_ctx *ctx, void *spec, int dtype, int loc);
/* ---- learned column physics (no counterpart in the reference) --------------
 * A column MLP stands in for the simple-physics column call of
 * phy_run_simple.f90:66-67: inputs T, q, u, v (nlev each, top -> bottom), ps,
 * sea-surface temperature, sin(lat); outputs the step's changes of Tv, q, u,
 * v; GELU hidden layers; bf16 weights on the matrix cores with FP32
 * accumulation, FP32 master weights, AdamW. The plain physics of the context
 * (physics = GFSDC_PHYS_SIMPLE) generates the training data, judges every
 * judge_stride-th column each step and replaces the whole step's physics when
 * the judged rms error (in units of each output's standard deviation)
 * exceeds the threshold. */
int gfsdc_mlp_init(gfsdc_ctx *ctx, int hidden, int nlayers, int maxbatch, long capacity,
                   unsigned long long seed);
/* advance nsteps with the plain physics, keeping every stride-th column of
 * each step (random offset) as (features, targets); returns the total count */
int gfsdc_mlp_collect(gfsdc_ctx *ctx, int nsteps, int stride, long *ncollected);
/* per-feature mean / standard deviation of the collected data */
int gfsdc_mlp_normalize(gfsdc_ctx *ctx);
/* nsteps AdamW steps on random batches from the first ntrain collected rows
 * (0: all); losses (nsteps, may be NULL) = batch mean squared normalised error */
int gfsdc_mlp_train(gfsdc_ctx *ctx, int nsteps, int batch, double lr, double wd, long ntrain,
                    double *losses);
/* rms normalised error of the network on collected rows [first, first+count)
 * and of the mean predictor (no physics change beyond the mean) on them */
int gfsdc_mlp_eval(gfsdc_ctx *ctx, long first, long count, double *err, double *base);
/* use the network in advance (on != 0) with its judge and fallback */
int gfsdc_mlp_enable(gfsdc_ctx *ctx, int on, int judge_stride, double threshold);
/* judge statistics since enable: last error, fallback steps, judged steps */
int gfsdc_mlp_judge_stats(gfsdc_ctx *ctx, double *out3);
/* parameters in the layout W_1[nin][hidden], b_1[hidden], ..., W_L[hidden][nout], b_L[nout] */
long gfsdc_mlp_nparams(gfsdc_ctx *ctx);
int gfsdc_mlp_params(gfsdc_ctx *ctx, float *get, const float *set);
/* the input and output normalisation [mean_x | std_x | mean_y | std_y]
 * (2 (nin + nout) floats; gfsdc_mlp_normalize makes it from the collected
 * columns): with gfsdc_mlp_params it moves a trained network to another
 * context of the same levels (a network trained on T126L64 columns run at
 * T254L64). Setting it refuses non-finite means and non-positive spreads. */
int gfsdc_mlp_norm(gfsdc_ctx *ctx, float *get, const float *set);
/* network on n normalised host rows x[n][nin] -> y[n][nout] */
int gfsdc_mlp_forward(gfsdc_ctx *ctx, const float *x, long n, float *y);
/* mean-squared-error loss and its parameter gradient on n <= maxbatch normalised host rows */
int gfsdc_mlp_grad(gfsdc_ctx *ctx, const float *x, const float *t, int n, float *grad, double *loss);
/* one AdamW step (beta 0.9/0.999, eps 1e-8, decay on weights only) with a host
 * gradient in the parameter layout, or the last gfsdc_mlp_grad when grad is NULL */
int gfsdc_mlp_adamw(gfsdc_ctx *ctx, const float *grad, double lr, double wd);

/* grid fields after the last getdyntend, names of grid_data.f90/pressure_data.f90:
 * 0 ug,1 vg,2 vrtg,3 divg,4 virtempg,5 dlnpdtg,6 etadot (nlev+1),7 psg,8 dlnpsdt,
 * 9 prs,14 dlnpsdx,15 dlnpsdy, 16 tracerg (nlev*ntrac), 17 precip (mm/s),
 * 18 pwat (kg/m^2) (phy_data_simple.f90:13; 2-D fields have one level); per member */
int gfsdc_get_grid(gfsdc_ctx *ctx, int which, void *out, int dtype, int loc);
/* semimp_data.f90 operators (FP64 host arrays, column-major as the reference):
 * amhyb,bmhyb (nlev,nlev), svhyb,tor_hyb (nlev), d_hyb_m (nlev,nlev,ntrunc+1,3) */
int gfsdc_get_semimp(gfsdc_ctx *ctx, double *amhyb, double *bmhyb, double *svhyb,
                     double *tor_hyb, double *d_hyb_m);
/* dyn_init.f90:450 setdampspec outputs (FP64 host): disspec (ndimspec), diff_prof, dmp_prof (nlev) */
int gfsdc_get_damp(gfsdc_ctx *ctx, double *disspec, double *diff_prof, double *dmp_prof);
/* run.f90:159-162 per-step diagnostics per member: spdmax, min ps, max ps (Pa), pstend */
int gfsdc_step_diag(gfsdc_ctx *ctx, double *out /* 4 * nmem */);
/* synthetic state: resting isothermal atmosphere (t0 K, ps0 Pa) plus seeded
 * random vorticity in total wavenumbers [nmin, nmax] scaled to an rms grid
 * vorticity of vort_rms (s^-1), member i seeded with seed + i */
int gfsdc_init_rest_state(gfsdc_ctx *ctx, double t0, double ps0, double vort_rms, int nmin,
                          int nmax, unsigned long long seed);
/* stochastic physics, the params.f90:140-160 namelist entries (an amplitude of
 * 0 switches a scheme off): vorticity confinement vcamp (m/s) modulated by
 * the svc pattern (dyn_run.f90:84-117,230-248); the sppt and shum patterns of
 * the GFS physics (their grids are made each step, run.f90:277-286, and handed
 * out by gfsdc_get_stoch; the GFS physics that applies them is not part of
 * this library); additive streamfunction noise addnoise per level
 * (stoch_data.f90:254-354, run.f90:288-292,514-518,541-552; the balanced
 * temperature part needs addnoise_vrtonly = 0 and a GSI balance file, and is
 * refused). Patterns are drawn from the counter RNG with seed iseed_* + member
 * (the reference seeds a Mersenne twister, 0 meaning the clock). */
typedef struct {
    double vcamp;
    double svc, svc_tau, svc_lscale;
    int iseed_svc;
    double sppt, sppt_tau, sppt_lscale;
    int iseed_sppt;
    double shum, shum_tau, shum_lscale;
    int iseed_shum;
    double addnoise, addnoise_tau, addnoise_lscale;
    int iseed_addnoise;
    int addnoise_vfilt, addnoise_kenorm, addnoise_dissfact, addnoise_vrtonly;
} gfsdc_stoch_cfg;
/* params.f90 defaults: everything off, addnoise_vrtonly = 1 */
void gfsdc_stoch_defaults(gfsdc_stoch_cfg *s);
/* replaces stoch_data.f90:37 init_stochdata: sets up and spins up the patterns
 * (5 e-folding times) of every scheme switched on; may be called again */
int gfsdc_stoch_init(gfsdc_ctx *ctx, const gfsdc_stoch_cfg *s);
/* which: 0..2 grids of svc, sppt (after its logit transform), shum
 * (nlon, nlat, nmem); 3..5 their spectral patterns (ndimspec, nmem); 6 the
 * addnoise patterns (ndimspec, nlev, nmem); 7 the last step's streamfunction
 * forcing psiforcing (ndimspec, nlev, nmem) */
int gfsdc_get_stoch(gfsdc_ctx *ctx, int which, void *out, int dtype, int loc);
/* model clock (s) at the start of the next step: run.f90:53 t = tstart
 * (params.f90: the initial file's fhour * 3600), advanced by dt per step
 * (run.f90:153); it times the IAU forcing */
int gfsdc_set_time(gfsdc_ctx *ctx, double t);
int gfsdc_get_time(gfsdc_ctx *ctx, double *t);
/* replaces iau_module.f90:20 init_iau (the increments analysis - first guess,
 * formed by the caller from the sigma files) and the getiauforcing calls of
 * run.f90:250-258: nfiles >= 2 increment times fhrs (hours, evenly spaced),
 * scaled by 1 / (delthrs h); each field in the set_state layout with a
 * trailing nfiles dimension. nfiles = 0 switches the forcing off. */
int gfsdc_iau_set(gfsdc_ctx *ctx, int nfiles, const double *fhrs, double delthrs, const void *vrt,
                  const void *div, const void *tv, const void *trac, const void *lnps, int dtype,
                  int loc);
/* replaces run.f90:58-145 (digital filter loop with set_dfi_wts :567-587):
 * 2 ntdfi steps, the state replaced by the Lanczos-weighted average of the
 * window, the clock set to the window's middle (tstart + ntdfi dt) */
int gfsdc_dfi(gfsdc_ctx *ctx, int ntdfi, int use_graph);
/* device checkpoint / resume (the restart that wrtout_sig / readin_sig give the
 * reference, dyn_init.f90:120-191, without the file format's FP32 rounding):
 * spectral state, orography, pdryini, clock, stochastic pattern and counters
 * in one image of gfsdc_checkpoint_size() bytes, at loc (host or device).
 * Loading checks that the image came from a context of the same precision,
 * sizes and pattern setting; continuing from it is bit-identical. */
int gfsdc_checkpoint_size(gfsdc_ctx *ctx, size_t *nbytes);
int gfsdc_checkpoint_save(gfsdc_ctx *ctx, void *buf, size_t nbytes, int loc);
int gfsdc_checkpoint_load(gfsdc_ctx *ctx, const void *buf, size_t nbytes, int loc);
/* wait for all work on the context's stream */
int gfsdc_sync(gfsdc_ctx *ctx);
/* HIP stream the context launches on (hipStream_t as void*) */
void *gfsdc_stream(gfsdc_ctx *ctx);

/* per-kernel timing of one step via HIP events on the context stream:
 * fills up to nmax (name, milliseconds) pairs, returns the count. The step is
 * run as gfsdc_advance runs all but its last, so after a lean Held-Suarez
 * step gfsdc_get_grid and gfsdc_step_diag return GFSDC_ESTATE until the next
 * advance or getdyntend. */
int gfsdc_profile_step(gfsdc_ctx *ctx, char names[][48], double *ms, int nmax);

#ifdef __cplusplus
}
#endif
#endif
