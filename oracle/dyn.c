/*
 * FP64 CPU restatement of the GFS hybrid sigma-pressure dynamical core step
 * (TEST INFRASTRUCTURE ONLY -- see oracle.h). Each function names the
 * reference lines it follows. Index comments use the reference's 1-based
 * level numbering: L(X,k) is level k of a 3-D grid array.
 *
 * Single-precision literals: the reference compiled with default_real=2
 * (kinds.f90:100) still writes constants as default-real literals
 * (physcons.f90:17-56, semimp_data.f90:74-88, phy_run_hs.f90:42-49,
 * pressure_data.f90:92-93). They are rounded to float here with SGL() so the
 * restatement carries the same values as that build.
 *
 * Pinning (tests/test_oracle_*.py):
 *  - calc_pressdata against closed forms;
 *  - getomega on general columns (non-isothermal, level-dependent winds,
 *    V.grad ln ps != 0, sigma and hybrid levels) by the column continuity
 *    equation, each layer's discrete mass budget, omega/p in the Simmons &
 *    Burridge (1981) form and their energy-conversion identity, all built in
 *    numpy with complex-step gradients;
 *  - getpresgrad on the same columns against -(grad phi + R T (grad ln p))
 *    with the hydrostatic geopotential, the sigma-coordinate textbook form,
 *    exact isothermal balance and the column angular-momentum identity;
 *  - energy conversion and continuity in getdyntend by the linearisation
 *    about a resting isothermal state against the independently built
 *    init_semimpdata operators, and by an exact balanced zonal flow;
 *  - init_semimpdata by (I + c^2 dt^2 n(n+1) ym) d = I and the tableau
 *    identities; setdampspec (sponge included) by its closed form;
 *  - advance by a generic IMEX-RK step (dense implicit solve per
 *    coefficient, backward-Euler damping) from the oracle's own tendencies,
 *    and by second-order convergence to expm(J t) on linear gravity waves;
 *  - HS by closed-form columns.
 * The vertical differencing of getvadv has no closed form and no
 * reference-run vectors here: it is pinned by the constant-field and
 * isothermal properties and by its identity with the flux-form scheme of
 * getvadv_tracers on linear profiles (limiter inactive), in every layer.
 * getvadv_tracers is pinned by its invariants: the flux part telescopes, a
 * uniform field has no tendency, a linear profile gives the centred getvadv
 * difference, and a forward step keeps a tracer non-negative.
 */
#include "oracle.h"
#include <float.h>
#include <math.h>
#include <omp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif
#define SGL(x) ((double)(float)(x))

/* IMEX tableau, semimp_data.f90:74-88 (default-real literals) */
#define A21 SGL(1. / 3.)
#define A31 SGL(1. / 6.)
#define A32 SGL(1. / 2.)
#define AA21 SGL(1. / 6.)
#define AA22 SGL(1. / 6.)
#define AA31 SGL(1. / 3.)
#define AA32 0.0
#define AA33 SGL(1. / 3.)
#define B1 SGL(1. / 2.)
#define B2 SGL(-1. / 2.)
#define B3 1.0
#define BB1 SGL(3. / 8.)
#define BB2 0.0
#define BB3 SGL(3. / 8.)
#define BB4 SGL(1. / 4.)

/* one AR(1) pattern kind: nf independent 2-D patterns (nf x nlm), noise drawn
 * from stream sbase + ctr nf + field (ctr counts advances) */
struct orc_pat {
    int nf;
    ocplx *spec;
    double *var, *vfact, phi, stdev;
    uint64_t seed, sbase, ctr;
};

struct orc_model {
    orc_cfg c;
    orc_sht *sht;
    int nlon, nlat, nlev, ntrac, T, nlm;
    long ng;
    double *ak, *bk, *ck, *dbk, *si, *sl;
    double *lats, *areawts, *lap, *sinlat, *coslat;
    int *degree;
    /* spectral state */
    ocplx *vrt, *div, *tv, *trac, *lnps, *topo;
    /* grid fields */
    double *ug, *vg, *vrtg, *divg, *tvg, *tracg, *dlnpdtg, *etadot, *lnpsg, *psg, *dlnpsdt;
    double *dphisdx, *dphisdy, *dlnpsdx, *dlnpsdy;
    double *precip, *pwat; /* simple physics: precipitation rate (mm/s), column water (kg/m^2) */
    double *pk, *dpk, *alfa, *rlnp, *prs;
    double *w[10]; /* work grids (nlev) */
    double *cb, *db, *cg, *wk1;
    /* semi-implicit and damping operators */
    double *amhyb, *bmhyb, *svhyb, *tor_hyb, *d_hyb_m;
    double *disspec, *diff_prof, *dmp_prof;
    double pdryini, kappa;
    ocplx *lin_d[3], *lin_t[3], *lin_p[3]; /* linear terms of the RK stages */
    int have_grid;
    /* AR(1) temperature pattern */
    ocplx *tpat;
    double *tpat_var, *tpat_vfact, tpat_phi;
    uint64_t tpat_step;
    int *order;
    /* model clock (s) at the start of the next step (run.f90:53,153) */
    double time;
    /* incremental analysis update (iau_module.f90): nfiles increments of the
     * whole state, at iaufhrs hours, scaled by 1 / (iau_delthrs h) */
    int iau_n;
    double *iau_hrs, iau_dt;
    ocplx *iau_inc; /* per file: vrt, div, tv (nl nlm), trac (ntrac nl nlm), lnps (nlm) */
    /* stochastic physics, kinds 0 svc, 1 sppt, 2 shum, 3 addnoise */
    orc_stoch_cfg sc;
    int stoch_on[4], vc_on;
    struct orc_pat pat[4];
    double *pgrd[3];    /* grids of svc, sppt (logit transform), shum */
    double *vcx, *vcy;  /* vorticity gradient (nlev grids) */
    ocplx *psif;        /* additive streamfunction forcing (nlev x nlm) */
    double *smoothfact; /* nlm */
};

#define L(X, k) ((X) + (long)((k) - 1) * m->ng)

void orc_cfg_defaults(orc_cfg *c) {
    memset(c, 0, sizeof(*c));
    c->explicit_ = 0;
    c->physics = 1;
    c->massfix = 1;
    c->ndiss = 0;
    c->efold = 0.0;
    c->hdif_fac = 1.0;
    c->hdif_fac2 = 1.0;
    c->fshk = 0.0;
    c->taustratdamp = SGL(5. * 86400.);
    /* physcons.f90:23-36 */
    c->rerth = SGL(6.3712e+6);
    c->grav = SGL(9.80665e+0);
    c->omega = SGL(7.2921e-5);
    c->rd = SGL(2.8705e+2);
    c->rv = SGL(4.6150e+2);
    c->cp = SGL(1.0046e+3);
    c->cvap = SGL(1.8460e+3);
    c->simple_test = 0;
    c->turbflux = 1;
}

/* OpenMP threads of the oracle's parallel loops (CPU-baseline timing) */
void orc_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }
int orc_get_threads(void) { return omp_get_max_threads(); }

/* counter-based RNG: splitmix64 finaliser over a mixed (seed, stream, index) key */
uint64_t orc_rng_u64(uint64_t seed, uint64_t stream, uint64_t index) {
    uint64_t z = seed * 0x9E3779B97F4A7C15ULL ^ (stream + 0x632BE59BD9B4E019ULL) * 0xBF58476D1CE4E5B9ULL;
    z += index * 0x94D049BB133111EBULL + 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
double orc_rng_uniform(uint64_t seed, uint64_t stream, uint64_t index) {
    return (double)(orc_rng_u64(seed, stream, index) >> 11) * 0x1p-53;
}

/* ---- pressure_data.f90:58-100 calc_pressdata ---------------------- */
static void calc_press_cols(int nlev, long ncol, const double *ak, const double *bk, double rk,
                            const double *lnpsg, double *psg, double *pk, double *dpk,
                            double *alfa, double *rlnp, double *prs) {
#pragma omp parallel for
    for (long i = 0; i < ncol; i++) {
        double ps = exp(lnpsg[i]);
        if (psg) psg[i] = ps;
        for (int k = 1; k <= nlev + 1; k++) pk[(long)(k - 1) * ncol + i] = ak[k - 1] + bk[k - 1] * ps;
        for (int k = 1; k <= nlev; k++) {
            double p1 = pk[(long)(k - 1) * ncol + i], p2 = pk[(long)k * ncol + i];
            double d = p2 - p1;
            dpk[(long)(k - 1) * ncol + i] = d;
            prs[(long)(nlev - k) * ncol + i] =
                pow((pow(p2, rk) * p2 - pow(p1, rk) * p1) / ((rk + 1.0) * d), 1.0 / rk);
        }
        alfa[i] = SGL(log(2.));
        rlnp[i] = SGL(99999.99);
        for (int k = 2; k <= nlev; k++) {
            double p1 = pk[(long)(k - 1) * ncol + i], p2 = pk[(long)k * ncol + i];
            double r = log(p2 / p1);
            rlnp[(long)(k - 1) * ncol + i] = r;
            alfa[(long)(k - 1) * ncol + i] = 1.0 - (p1 / dpk[(long)(k - 1) * ncol + i]) * r;
        }
    }
}

void orc_calc_pressdata(int nlev, int ncol, const double *ak, const double *bk, double rk,
                        const double *lnpsg, double *pk, double *dpk, double *alfa,
                        double *rlnp, double *prs) {
    calc_press_cols(nlev, ncol, ak, bk, rk, lnpsg, NULL, pk, dpk, alfa, rlnp, prs);
}

/* ---- semimp_data.f90:127-256 init_semimpdata ---------------------- */
static int invert(int n, double *a /* col-major n*n, replaced by inverse */) {
    double *b = calloc((size_t)n * n, sizeof(double));
    for (int i = 0; i < n; i++) b[i + i * n] = 1.0;
    for (int c = 0; c < n; c++) {
        int piv = c;
        for (int r = c + 1; r < n; r++)
            if (fabs(a[r + c * n]) > fabs(a[piv + c * n])) piv = r;
        if (a[piv + c * n] == 0.0) { free(b); return -1; }
        if (piv != c)
            for (int j = 0; j < n; j++) {
                double t = a[c + j * n]; a[c + j * n] = a[piv + j * n]; a[piv + j * n] = t;
                t = b[c + j * n]; b[c + j * n] = b[piv + j * n]; b[piv + j * n] = t;
            }
        double d = 1.0 / a[c + c * n];
        for (int j = 0; j < n; j++) { a[c + j * n] *= d; b[c + j * n] *= d; }
        for (int r = 0; r < n; r++) {
            if (r == c) continue;
            double f = a[r + c * n];
            if (f == 0.0) continue;
            for (int j = 0; j < n; j++) { a[r + j * n] -= f * a[c + j * n]; b[r + j * n] -= f * b[c + j * n]; }
        }
    }
    memcpy(a, b, sizeof(double) * n * n);
    free(b);
    return 0;
}

static void init_semimp(orc_model *m) {
    int nl = m->nlev;
    double rd = m->c.rd, kappa = m->kappa, rerth = m->c.rerth, dt = m->c.dt;
    const double tref = 300.0, pref = 800.e2; /* semimp_data.f90:122-123 */
    double *pkref = malloc(sizeof(double) * (nl + 1)), *dpkref = malloc(sizeof(double) * nl);
    double *alfaref = malloc(sizeof(double) * nl);
    double *yecm = calloc((size_t)nl * nl, sizeof(double)), *tecm = calloc((size_t)nl * nl, sizeof(double));
    double *ym = calloc((size_t)nl * nl, sizeof(double));
#define CM(A, i, j) A[((i) - 1) + (long)((j) - 1) * nl]
    for (int k = 1; k <= nl + 1; k++) pkref[k - 1] = m->ak[k - 1] + m->bk[k - 1] * pref;
    for (int k = 1; k <= nl; k++) dpkref[k - 1] = pkref[k] - pkref[k - 1];
    alfaref[0] = SGL(log(2.));
    for (int k = 2; k <= nl; k++)
        alfaref[k - 1] = 1.0 - (pkref[k - 1] / dpkref[k - 1]) * log(pkref[k] / pkref[k - 1]);
    for (int ir = 1; ir <= nl; ir++) {
        CM(yecm, ir, ir) = alfaref[ir - 1] * rd;
        for (int ic = ir + 1; ic <= nl; ic++) CM(yecm, ir, ic) = rd * log(pkref[ic] / pkref[ic - 1]);
    }
    for (int ir = 1; ir <= nl; ir++) {
        CM(tecm, ir, ir) = kappa * tref * alfaref[ir - 1];
        for (int ic = 1; ic <= ir - 1; ic++)
            CM(tecm, ir, ic) = (kappa * tref * dpkref[ic - 1] / dpkref[ir - 1]) * log(pkref[ir] / pkref[ir - 1]);
    }
    for (int j = 1; j <= nl; j++) {
        m->svhyb[j - 1] = dpkref[nl - j] / pref;
        for (int k = 1; k <= nl; k++) {
            CM(m->amhyb, k, j) = CM(yecm, nl + 1 - k, nl + 1 - j) / (rerth * rerth);
            CM(m->bmhyb, k, j) = CM(tecm, nl + 1 - k, nl + 1 - j);
        }
    }
    for (int k = 0; k < nl; k++) m->tor_hyb[k] = rd * tref / (rerth * rerth);
    for (int i = 1; i <= nl; i++)
        for (int j = 1; j <= nl; j++) {
            double s = m->tor_hyb[i - 1] * m->svhyb[j - 1];
            for (int k = 1; k <= nl; k++) s += CM(m->amhyb, i, k) * CM(m->bmhyb, k, j);
            CM(ym, i, j) = s;
        }
    const double consts[3] = {AA22, AA33, BB4};
    for (int s = 0; s < 3; s++)
        for (int nn = 1; nn <= m->T + 1; nn++) {
            int n = nn - 1;
            double rnn1 = (double)n * (n + 1);
            double *d = m->d_hyb_m + ((long)s * (m->T + 1) + (nn - 1)) * nl * nl;
            double cdt = consts[s] * dt;
            for (int i = 1; i <= nl; i++)
                for (int j = 1; j <= nl; j++) CM(d, i, j) = (i == j ? 1.0 : 0.0) + cdt * cdt * rnn1 * CM(ym, i, j);
            if (invert(nl, d) != 0) fprintf(stderr, "oracle: singular semi-implicit matrix\n");
        }
#undef CM
    free(pkref); free(dpkref); free(alfaref); free(yecm); free(tecm); free(ym);
}

/* ---- dyn_init.f90:450-494 setdampspec ------------------------------ */
static void setdampspec(orc_model *m) {
    int ndiss = m->c.ndiss == 0 ? 8 : m->c.ndiss;
    double efold = m->c.efold, fshk = m->c.fshk;
    double rerth = m->c.rerth, hf2 = m->c.hdif_fac2, hf = m->c.hdif_fac;
    int T = m->T;
    if (efold <= 0) {
        if (T > 170)
            efold = 3600. / (hf2 * SGL(pow((float)T / 170.f, 4)) * SGL(1.1));
        else if (T == 126 || T == 170)
            efold = 1. / (hf2 * SGL(12.E15) / pow(rerth, 4) * SGL(80. * 81.) * SGL(80. * 81.));
        else
            efold = 1. / (hf2 * SGL(3.E15) / pow(rerth, 4) * SGL(80. * 81.) * SGL(80. * 81.));
        efold = 2. * efold;
    }
    if (fshk <= 0) {
        fshk = 1.0 * hf;
        if (T > 170) fshk = SGL(2.2) * hf;
        if (T == 126) fshk = SGL(1.5) * hf;
    }
    double slrd0 = SGL(0.002), dmp1 = 1. / m->c.taustratdamp;
    for (int k = 0; k < m->nlev; k++) {
        m->dmp_prof[k] = 0.0;
        m->diff_prof[k] = pow(m->sl[k], log(1. / fshk));
        if (m->sl[k] < slrd0) m->dmp_prof[k] = dmp1 * log(slrd0 / m->sl[k]);
    }
    double minlap = -(double)T * (T + 1);
    for (int k = 0; k < m->nlm; k++) {
        double r = m->lap[k] / minlap, p = 1.0;
        for (int i = 0; i < ndiss / 2; i++) p *= r;
        m->disspec[k] = -(1. / efold) * p;
    }
}

/* ---- patterngenerator.f90:33-177 AR(1) spectral pattern --------------
 * Gaussian noise from the counter RNG (Box-Muller on two uniforms per
 * coefficient) in place of the reference's Mersenne twister stream:
 * member seed = tpat_seed (+ member), stream = 2^40 + advance count,
 * index = 2k, 2k+1. */
static void pat_noise(const orc_model *m, uint64_t seed, uint64_t stream, ocplx *noise) {
    const double normfact = 2.0 * M_PI; /* stoch_data.f90 passes 2 pi */
    const double sc = sqrt(normfact / m->T);
    for (int k = 0; k < m->nlm; k++) {
        if (k == 0) { noise[0] = 0.0; continue; } /* no global mean (115-116) */
        double u1 = orc_rng_uniform(seed, stream, 2 * (uint64_t)k);
        double u2 = orc_rng_uniform(seed, stream, 2 * (uint64_t)k + 1);
        double r = sqrt(-2.0 * log(1.0 - u1));
        double gr = r * cos(2.0 * M_PI * u2), gi = r * sin(2.0 * M_PI * u2);
        double n = m->degree[k], w = 1.0 / sqrt(2.0 * n + 1.0);
        noise[k] = (m->order[k] != 0) ? (gr + I * gi) * w * sc : sqrt(2.0) * gr * w * sc;
    }
}

static void pattern_noise(const orc_model *m, uint64_t step, ocplx *noise) {
    pat_noise(m, (uint64_t)(int64_t)m->c.tpat_seed, (1ULL << 40) + step, noise);
}

/* gaussian_spect (145-177): sqrt(T exp(L^2 lap / 4a^2)), normalised so that
 * the deterministic reference noise has unit grid variance */
static void pat_varspec(const orc_model *m, double L, double *var) {
    double a = m->c.rerth, v = 0.0;
    int T = m->T;
    for (int k = 0; k < m->nlm; k++) {
        double n = m->degree[k], lap = -n * (n + 1.0);
        var[k] = sqrt(T * exp(L * L * lap / (4.0 * a * a)));
        double nz2 = (k == 0) ? 0.0 : 2.0 / (2.0 * n + 1.0); /* |(1+i)|^2 or |sqrt 2|^2 over 2n+1 */
        nz2 *= (2.0 * M_PI) / T;
        double t = var[k] * var[k] * nz2;
        v += (m->order[k] != 0) ? t : 0.5 * t;
    }
    v /= 2.0 * M_PI;
    for (int k = 0; k < m->nlm; k++) var[k] /= sqrt(v);
}

void orc_pattern_advance(orc_model *m) {
    /* patterngenerator.f90:134-143 */
    ocplx *noise = malloc(sizeof(ocplx) * m->nlm);
    pattern_noise(m, m->tpat_step++, noise);
    double f = m->c.tpat_amp * sqrt(1.0 - m->tpat_phi * m->tpat_phi);
    for (int k = 0; k < m->nlm; k++) m->tpat[k] = m->tpat_phi * m->tpat[k] + f * m->tpat_var[k] * noise[k];
    free(noise);
}

static void pattern_init(orc_model *m) {
    int nlm = m->nlm;
    pat_varspec(m, m->c.tpat_lscale, m->tpat_var);
    m->tpat_phi = exp(-m->c.dt / m->c.tpat_tau);
    /* vertical taper of stoch_data.f90 (sigma 0.05 - 0.10) */
    for (int k = 0; k < m->nlev; k++) {
        double sl = m->sl[k];
        m->tpat_vfact[k] = 1.0;
        if (sl < 0.10 && sl > 0.05) m->tpat_vfact[k] = (sl - 0.05) / (0.10 - 0.05);
        else if (sl <= 0.05) m->tpat_vfact[k] = 0.0;
    }
    /* init_stochdata: start from scaled noise, spin up 5 e-folding times */
    m->tpat_step = 0;
    pattern_noise(m, m->tpat_step++, m->tpat);
    for (int k = 0; k < nlm; k++) m->tpat[k] *= m->c.tpat_amp * m->tpat_var[k];
    int nspinup = (int)(5.0 * m->c.tpat_tau / m->c.dt);
    for (int i = 0; i < nspinup; i++) orc_pattern_advance(m);
}

/* ---- stochastic physics (stoch_data.f90, run.f90:272-292,509-553) ------
 * Kinds 0 svc, 1 sppt, 2 shum (one 2-D pattern each) and 3 addnoise (one per
 * level), each an AR(1) process of patterngenerator.f90 with its own seed and
 * random stream (2^40 (3 + kind) + draw). */
static void pat_advance(orc_model *m, struct orc_pat *p, double phi, double f) {
    ocplx *noise = malloc(sizeof(ocplx) * m->nlm);
    for (int fld = 0; fld < p->nf; fld++) {
        pat_noise(m, p->seed, p->sbase + p->ctr * p->nf + fld, noise);
        ocplx *q = p->spec + (long)fld * m->nlm;
        for (int k = 0; k < m->nlm; k++) q[k] = phi * q[k] + f * p->var[k] * noise[k];
    }
    p->ctr++;
    free(noise);
}

/* vertical factors of stoch_data.f90: a taper to 0 between sigma 0.10 and
 * 0.05 (svc :63-71, sppt :85-93, addnoise :128-136), exp((sl - 1) / 0.2) for
 * shum (:103-106) */
static void stoch_vfact(const orc_model *m, int kind, double *vf) {
    for (int k = 0; k < m->nlev; k++) {
        double sl = m->sl[k];
        if (kind == 2) {
            vf[k] = exp((sl - 1.) / SGL(0.2));
            continue;
        }
        vf[k] = 1.0;
        if (sl < SGL(0.10) && sl > SGL(0.05)) vf[k] = (sl - SGL(0.05)) / (SGL(0.10) - SGL(0.05));
        else if (sl <= SGL(0.05)) vf[k] = 0.0;
    }
}

int orc_stoch_init(orc_model *m, const orc_stoch_cfg *s) {
    m->sc = *s;
    const double amp[4] = {s->svc, s->sppt, s->shum, s->addnoise};
    const double tau[4] = {s->svc_tau, s->sppt_tau, s->shum_tau, s->addnoise_tau};
    const double lsc[4] = {s->svc_lscale, s->sppt_lscale, s->shum_lscale, s->addnoise_lscale};
    const int seed[4] = {s->iseed_svc, s->iseed_sppt, s->iseed_shum, s->iseed_addnoise};
    int nlm = m->nlm, nl = m->nlev;
    m->vc_on = s->vcamp > FLT_MIN || s->svc > FLT_MIN;
    if (m->vc_on && !m->vcx) {
        m->vcx = malloc(sizeof(double) * m->ng * nl);
        m->vcy = malloc(sizeof(double) * m->ng * nl);
    }
    for (int kd = 0; kd < 4; kd++) {
        struct orc_pat *p = &m->pat[kd];
        m->stoch_on[kd] = amp[kd] > FLT_MIN;
        if (!m->stoch_on[kd]) continue;
        if (!(tau[kd] > 0) || (kd == 3 && !s->addnoise_vrtonly)) return -1;
        p->nf = kd == 3 ? nl : 1;
        free(p->spec); free(p->var); free(p->vfact);
        p->spec = calloc((size_t)p->nf * nlm, sizeof(ocplx));
        p->var = malloc(sizeof(double) * nlm);
        p->vfact = malloc(sizeof(double) * nl);
        pat_varspec(m, lsc[kd], p->var);
        stoch_vfact(m, kd, p->vfact);
        p->phi = exp(-m->c.dt / tau[kd]);
        p->stdev = amp[kd];
        p->seed = (uint64_t)(int64_t)seed[kd];
        p->sbase = (uint64_t)(3 + kd) << 40;
        p->ctr = 0;
        /* stdev x varspectrum x noise, then 5 e-folding times of spin-up */
        pat_advance(m, p, 0.0, p->stdev);
        int nspinup = (int)(5.0 * tau[kd] / m->c.dt);
        for (int i = 0; i < nspinup; i++) pat_advance(m, p, p->phi, p->stdev * sqrt(1.0 - p->phi * p->phi));
        if (kd < 3 && !m->pgrd[kd]) m->pgrd[kd] = calloc(m->ng, sizeof(double));
        /* the grid of the spun-up pattern, for tendencies asked before the first step */
        if (kd < 3) orc_spectogrd(m->sht, p->spec, m->pgrd[kd], 1);
        if (kd == 1)
            for (long i = 0; i < m->ng; i++) m->pgrd[1][i] = 2. / (1. + exp(m->pgrd[1][i])) - 1.;
    }
    if (m->stoch_on[3]) {
        if (!m->psif) m->psif = calloc((size_t)nl * nlm, sizeof(ocplx));
        /* getstochforcing :269-276: Gaussian smoothing at three times the noise scale */
        if (!m->smoothfact) m->smoothfact = malloc(sizeof(double) * nlm);
        double rnn1 = 2. * m->c.rerth / (3. * s->addnoise_lscale), rnn0 = rnn1 * (rnn1 + 1.);
        for (int k = 0; k < nlm; k++) {
            double n = m->degree[k];
            m->smoothfact[k] = exp(-(n * (n + 1.) / rnn0));
        }
    }
    return 0;
}

/* run.f90:272-292: the patterns used this step and their grids; sppt through
 * the logit transform 2 / (1 + e^g) - 1 */
static void stoch_patterns(orc_model *m) {
    for (int kd = 0; kd < 4; kd++) {
        struct orc_pat *p = &m->pat[kd];
        if (!m->stoch_on[kd]) continue;
        pat_advance(m, p, p->phi, p->stdev * sqrt(1.0 - p->phi * p->phi));
        if (kd < 3) orc_spectogrd(m->sht, p->spec, m->pgrd[kd], 1);
        if (kd == 1)
            for (long i = 0; i < m->ng; i++) m->pgrd[1][i] = 2. / (1. + exp(m->pgrd[1][i])) - 1.;
    }
}

/* getstochforcing (stoch_data.f90:254-354) with addnoise_vrtonly: the
 * streamfunction noise, optionally 1-2-1 filtered in the vertical (one pass
 * whatever addnoise_vfilt is, as the reference's loop re-filters the
 * unfiltered patterns), on the grid times vfact and, with addnoise_dissfact,
 * times the smoothed square root of the kinetic energy dissipation rate
 * |u . u_diss| (u_diss the wind of -disspec vrt, -disspec div), analysed back.
 * Without addnoise_dissfact the reference reads an array it never set; here
 * the forcing is then the filtered noise times vfact. On the grid fields of the
 * inverse transform before the physics. */
static void stoch_forcing(orc_model *m) {
    int nl = m->nlev, nlm = m->nlm;
    long ng = m->ng;
    const struct orc_pat *p = &m->pat[3];
    const double epstiny = FLT_MIN;
    ocplx *ps = m->psif;
    if (m->sc.addnoise_vfilt > 0) {
        const double c = SGL(1. + 1. / 3.);
        for (int k = 0; k < nl; k++) {
            const ocplx *q = p->spec + (long)k * nlm;
            for (int n = 0; n < nlm; n++) {
                ocplx w;
                if (k == 0)
                    w = c * q[nlm + n] + 2. * c * q[n];
                else if (k == nl - 1)
                    w = c * q[n - nlm] + 2. * c * q[n];
                else
                    w = q[n + nlm] + 2. * q[n] + q[n - nlm];
                ps[(long)k * nlm + n] = 0.25 * w;
            }
        }
    } else {
        memcpy(ps, p->spec, sizeof(ocplx) * nl * nlm);
    }
    double *ud = malloc(sizeof(double) * ng), *vd = malloc(sizeof(double) * ng);
    double *wg = malloc(sizeof(double) * ng);
    ocplx *w1 = malloc(sizeof(ocplx) * nlm), *w2 = malloc(sizeof(ocplx) * nlm);
    for (int k = 0; k < nl; k++) {
        const long o = (long)k * ng;
        orc_spectogrd(m->sht, ps + (long)k * nlm, wg, 1);
        if (m->sc.addnoise_dissfact) {
            for (int n = 0; n < nlm; n++) {
                w1[n] = -m->disspec[n] * m->vrt[(long)k * nlm + n];
                w2[n] = -m->disspec[n] * m->div[(long)k * nlm + n];
            }
            orc_getuv(m->sht, w1, w2, ud, vd, 1, m->c.rerth);
            for (long i = 0; i < ng; i++) {
                double d = fabs(m->ug[o + i] * ud[i] + m->vg[o + i] * vd[i]);
                ud[i] = sqrt(d < epstiny ? epstiny : d);
            }
            orc_grdtospec(m->sht, ud, w1, 1);
            for (int n = 0; n < nlm; n++) w1[n] *= m->smoothfact[n];
            orc_spectogrd(m->sht, w1, vd, 1);
            for (long i = 0; i < ng; i++) ud[i] = p->vfact[k] * wg[i] * vd[i];
        } else {
            for (long i = 0; i < ng; i++) ud[i] = p->vfact[k] * wg[i];
        }
        orc_grdtospec(m->sht, ud, ps + (long)k * nlm, 1);
    }
    free(ud); free(vd); free(wg); free(w1); free(w2);
}

/* run.f90:541-552: the streamfunction forcing into vorticity (lap / a^2, or
 * -sqrt(-lap) / a with addnoise_kenorm and addnoise_vrtonly); the balanced
 * temperature forcing is zero with addnoise_vrtonly */
static void stoch_apply(orc_model *m) {
    double dt = m->c.dt, a = m->c.rerth;
    int ke = m->sc.addnoise_kenorm && m->sc.addnoise_vrtonly;
    for (int k = 0; k < m->nlev; k++)
        for (int n = 0; n < m->nlm; n++) {
            long i = (long)k * m->nlm + n;
            if (ke)
                m->vrt[i] -= dt * (sqrt(-m->lap[n]) / a) * m->psif[i];
            else
                m->vrt[i] += dt * (m->lap[n] / (a * a)) * m->psif[i];
        }
}

/* getstochforcing on the current state and grid fields (after an
 * inverse-only getdyntend); the result in psiforcing */
void orc_stoch_forcing(orc_model *m) {
    if (m->stoch_on[3] && m->have_grid) stoch_forcing(m);
}

/* replace a spectral pattern (which 3..5: svc, sppt, shum; 6: addnoise) */
void orc_set_stoch(orc_model *m, int which, const void *in) {
    if (which >= 3 && which < 6 && m->stoch_on[which - 3])
        memcpy(m->pat[which - 3].spec, in, sizeof(ocplx) * m->nlm);
    else if (which == 6 && m->stoch_on[3])
        memcpy(m->pat[3].spec, in, sizeof(ocplx) * m->nlm * m->nlev);
}

void orc_get_stoch(const orc_model *m, int which, void *out) {
    if (which >= 0 && which < 3) {
        if (m->stoch_on[which]) memcpy(out, m->pgrd[which], sizeof(double) * m->ng);
    } else if (which >= 3 && which < 6) {
        if (m->stoch_on[which - 3]) memcpy(out, m->pat[which - 3].spec, sizeof(ocplx) * m->nlm);
    } else if (which == 6) {
        if (m->stoch_on[3]) memcpy(out, m->pat[3].spec, sizeof(ocplx) * m->nlm * m->nlev);
    } else if (which == 7) {
        if (m->stoch_on[3]) memcpy(out, m->psif, sizeof(ocplx) * m->nlm * m->nlev);
    } else if (which >= 8 && which < 12) {
        if (m->stoch_on[which - 8]) memcpy(out, m->pat[which - 8].vfact, sizeof(double) * m->nlev);
    }
}

void orc_get_pattern(const orc_model *m, ocplx *pattern, double *varspectrum, double *phi, double *vfact) {
    if (pattern) memcpy(pattern, m->tpat, sizeof(ocplx) * m->nlm);
    if (varspectrum) memcpy(varspectrum, m->tpat_var, sizeof(double) * m->nlm);
    if (phi) *phi = m->tpat_phi;
    if (vfact) memcpy(vfact, m->tpat_vfact, sizeof(double) * m->nlev);
}

orc_model *orc_model_create(const orc_cfg *c) {
    orc_model *m = calloc(1, sizeof(orc_model));
    m->c = *c;
    m->nlon = c->nlon; m->nlat = c->nlat; m->nlev = c->nlev; m->ntrac = c->ntrac; m->T = c->ntrunc;
    m->sht = orc_sht_create(c->nlon, c->nlat, c->ntrunc);
    m->nlm = orc_sht_nlm(m->sht);
    m->ng = (long)c->nlon * c->nlat;
    m->kappa = c->rd / c->cp;
    int nl = c->nlev, nlm = m->nlm;
    long ng = m->ng;
    m->ak = malloc(sizeof(double) * (nl + 1)); m->bk = malloc(sizeof(double) * (nl + 1));
    memcpy(m->ak, c->ak, sizeof(double) * (nl + 1));
    memcpy(m->bk, c->bk, sizeof(double) * (nl + 1));
    m->ck = malloc(sizeof(double) * nl); m->dbk = malloc(sizeof(double) * nl);
    m->si = malloc(sizeof(double) * (nl + 1)); m->sl = malloc(sizeof(double) * nl);
    for (int k = 1; k <= nl; k++) {
        m->dbk[k - 1] = m->bk[k] - m->bk[k - 1];
        m->ck[k - 1] = m->ak[k] * m->bk[k - 1] - m->ak[k - 1] * m->bk[k];
    }
    /* pressure_data.f90:63-70 */
    for (int k = 1; k <= nl + 1; k++) m->si[nl + 1 - k] = m->ak[k - 1] / SGL(101300.0) + m->bk[k - 1];
    for (int k = 0; k < nl; k++) m->sl[k] = 0.5 * (m->si[k] + m->si[k + 1]);
    m->lats = malloc(sizeof(double) * c->nlat);
    m->areawts = malloc(sizeof(double) * ng);
    m->lap = malloc(sizeof(double) * nlm);
    m->degree = malloc(sizeof(int) * nlm);
    orc_sht_griddata(m->sht, m->lats, NULL, m->areawts, m->degree, NULL, m->lap, NULL);
    m->sinlat = malloc(sizeof(double) * ng);
    m->coslat = malloc(sizeof(double) * ng);
    for (int j = 0; j < c->nlat; j++)
        for (int i = 0; i < c->nlon; i++) {
            m->sinlat[(long)j * c->nlon + i] = sin(m->lats[j]);
            m->coslat[(long)j * c->nlon + i] = cos(m->lats[j]);
        }
    size_t s3 = sizeof(ocplx) * nlm * nl;
    m->vrt = calloc(1, s3); m->div = calloc(1, s3); m->tv = calloc(1, s3);
    m->trac = calloc(1, s3 * (c->ntrac > 0 ? c->ntrac : 1));
    m->lnps = calloc(nlm, sizeof(ocplx)); m->topo = calloc(nlm, sizeof(ocplx));
    size_t g3 = sizeof(double) * ng * nl, g2 = sizeof(double) * ng;
    m->ug = malloc(g3); m->vg = malloc(g3); m->vrtg = malloc(g3); m->divg = malloc(g3);
    m->tvg = malloc(g3); m->tracg = malloc(g3 * (c->ntrac > 0 ? c->ntrac : 1));
    m->dlnpdtg = malloc(g3); m->etadot = malloc(g3 + g2); m->prs = malloc(g3);
    m->pk = malloc(g3 + g2); m->dpk = malloc(g3); m->alfa = malloc(g3); m->rlnp = malloc(g3);
    for (int i = 0; i < 10; i++) m->w[i] = malloc(g3);
    m->cb = malloc(g3); m->db = malloc(g3); m->cg = malloc(g3); m->wk1 = malloc(g3 + g2);
    m->lnpsg = malloc(g2); m->psg = malloc(g2); m->dlnpsdt = malloc(g2);
    m->dphisdx = calloc(ng, sizeof(double)); m->dphisdy = calloc(ng, sizeof(double));
    m->dlnpsdx = malloc(g2); m->dlnpsdy = malloc(g2);
    m->precip = calloc(ng, sizeof(double)); m->pwat = calloc(ng, sizeof(double));
    m->amhyb = malloc(sizeof(double) * nl * nl); m->bmhyb = malloc(sizeof(double) * nl * nl);
    m->svhyb = malloc(sizeof(double) * nl); m->tor_hyb = malloc(sizeof(double) * nl);
    m->d_hyb_m = malloc(sizeof(double) * nl * nl * (m->T + 1) * 3);
    m->disspec = malloc(sizeof(double) * nlm);
    m->diff_prof = malloc(sizeof(double) * nl); m->dmp_prof = malloc(sizeof(double) * nl);
    for (int s = 0; s < 3; s++) {
        m->lin_d[s] = malloc(s3); m->lin_t[s] = malloc(s3); m->lin_p[s] = malloc(sizeof(ocplx) * nlm);
    }
    init_semimp(m);
    setdampspec(m);
    m->order = malloc(sizeof(int) * nlm);
    orc_sht_griddata(m->sht, NULL, NULL, NULL, NULL, m->order, NULL, NULL);
    m->tpat = calloc(nlm, sizeof(ocplx));
    m->tpat_var = calloc(nlm, sizeof(double));
    m->tpat_vfact = calloc(nl, sizeof(double));
    if (c->tpat_amp > 0) pattern_init(m);
    return m;
}

void orc_model_destroy(orc_model *m) {
    if (!m) return;
    orc_sht_destroy(m->sht);
    void *ptrs[] = {m->ak, m->bk, m->ck, m->dbk, m->si, m->sl, m->lats, m->areawts, m->lap,
                    m->degree, m->sinlat, m->coslat, m->vrt, m->div, m->tv, m->trac, m->lnps,
                    m->topo, m->ug, m->vg, m->vrtg, m->divg, m->tvg, m->tracg, m->dlnpdtg,
                    m->etadot, m->prs, m->pk, m->dpk, m->alfa, m->rlnp, m->cb, m->db, m->cg,
                    m->wk1, m->lnpsg, m->psg, m->dlnpsdt, m->dphisdx, m->dphisdy, m->dlnpsdx,
                    m->dlnpsdy, m->amhyb, m->bmhyb, m->svhyb, m->tor_hyb, m->d_hyb_m,
                    m->disspec, m->diff_prof, m->dmp_prof, m->precip, m->pwat, m->order,
                    m->tpat, m->tpat_var, m->tpat_vfact, m->iau_hrs, m->iau_inc, m->pgrd[0],
                    m->pgrd[1], m->pgrd[2], m->vcx, m->vcy, m->psif, m->smoothfact,
                    m->pat[0].spec, m->pat[0].var, m->pat[0].vfact, m->pat[1].spec, m->pat[1].var,
                    m->pat[1].vfact, m->pat[2].spec, m->pat[2].var, m->pat[2].vfact, m->pat[3].spec,
                    m->pat[3].var, m->pat[3].vfact};
    for (size_t i = 0; i < sizeof(ptrs) / sizeof(ptrs[0]); i++) free(ptrs[i]);
    for (int i = 0; i < 10; i++) free(m->w[i]);
    for (int s = 0; s < 3; s++) { free(m->lin_d[s]); free(m->lin_t[s]); free(m->lin_p[s]); }
    free(m);
}

const orc_sht *orc_model_sht(const orc_model *m) { return m->sht; }

void orc_set_state(orc_model *m, const ocplx *vrt, const ocplx *div, const ocplx *tv,
                   const ocplx *trac, const ocplx *lnps) {
    size_t s3 = sizeof(ocplx) * m->nlm * m->nlev;
    if (vrt) memcpy(m->vrt, vrt, s3);
    if (div) memcpy(m->div, div, s3);
    if (tv) memcpy(m->tv, tv, s3);
    if (trac && m->ntrac > 0) memcpy(m->trac, trac, s3 * m->ntrac);
    if (lnps) memcpy(m->lnps, lnps, sizeof(ocplx) * m->nlm);
}

void orc_get_state(const orc_model *m, ocplx *vrt, ocplx *div, ocplx *tv, ocplx *trac,
                   ocplx *lnps) {
    size_t s3 = sizeof(ocplx) * m->nlm * m->nlev;
    if (vrt) memcpy(vrt, m->vrt, s3);
    if (div) memcpy(div, m->div, s3);
    if (tv) memcpy(tv, m->tv, s3);
    if (trac && m->ntrac > 0) memcpy(trac, m->trac, s3 * m->ntrac);
    if (lnps) memcpy(lnps, m->lnps, sizeof(ocplx) * m->nlm);
}

void orc_set_pdryini(orc_model *m, double p) { m->pdryini = p; }
double orc_get_pdryini(const orc_model *m) { return m->pdryini; }
void orc_set_time(orc_model *m, double t) { m->time = t; }

/* dyn_init.f90:84: the orography's geopotential gradient,
 * getgrad(grav * topospec, dphisdx, dphisdy, rerth) */
void orc_set_topo(orc_model *m, const ocplx *topo) {
    ocplx *gt = malloc(sizeof(ocplx) * m->nlm);
    for (int k = 0; k < m->nlm; k++) {
        m->topo[k] = topo[k];
        gt[k] = m->c.grav * topo[k];
    }
    orc_getgrad(m->sht, gt, m->dphisdx, m->dphisdy, 1, m->c.rerth);
    free(gt);
}
double orc_get_time(const orc_model *m) { return m->time; }

/* ---- iau_module.f90:20-125 ------------------------------------------------
 * init_iau keeps, per file time iaufhrs(n), the increment analysis - first
 * guess of every state field (:64-80); the files are 2 or more at a constant
 * spacing (:38-51). inc holds the nfiles increments one after the other, each
 * as vrt, div, tv (nl, nlm), trac (ntrac, nl, nlm), lnps (nlm). */
static long state_len(const orc_model *m) { return (long)m->nlm * (m->nlev * (3 + m->ntrac) + 1); }

int orc_iau_set(orc_model *m, int nfiles, const double *fhrs, double delthrs, const ocplx *inc) {
    free(m->iau_hrs);
    free(m->iau_inc);
    m->iau_hrs = NULL;
    m->iau_inc = NULL;
    m->iau_n = 0;
    if (nfiles == 0) return 0;
    if (nfiles < 2) return -1;
    for (int n = 1; n < nfiles; n++)
        if (fhrs[n] - fhrs[n - 1] != fhrs[1] - fhrs[0] || !(fhrs[1] > fhrs[0])) return -1;
    long len = state_len(m);
    m->iau_hrs = malloc(sizeof(double) * nfiles);
    m->iau_inc = malloc(sizeof(ocplx) * len * nfiles);
    memcpy(m->iau_hrs, fhrs, sizeof(double) * nfiles);
    memcpy(m->iau_inc, inc, sizeof(ocplx) * len * nfiles);
    m->iau_n = nfiles;
    m->iau_dt = delthrs * 3600.;
    return 0;
}

/* getiauforcing (:84-115) at model time t: zero outside [iaufhrs(1),
 * iaufhrs(nfiles)] h, the first or last increment / dt at those end points,
 * else linear in time between the two files around t, ((1 - delt) inc(n) +
 * delt inc(n-1)) / dt with n the first file later than t. Returns 0 when the
 * forcing is zero. */
static int iau_forcing(const orc_model *m, double t, ocplx *F) {
    int nf = m->iau_n;
    long len = state_len(m);
    const double *h = m->iau_hrs;
    if (nf == 0 || t < h[0] * 3600. || t > h[nf - 1] * 3600.) return 0;
    double whi = 1.0, wlo = 0.0;
    int n;
    if (t == 3600. * h[nf - 1]) {
        n = nf - 1;
    } else if (t == 3600. * h[0]) {
        n = 0;
    } else {
        for (n = 0; n < nf; n++)
            if (h[n] * 3600. > t) break;
        double delt = (h[n] - t / 3600.) / (h[n] - h[n - 1]);
        whi = 1. - delt;
        wlo = delt;
    }
    const ocplx *hi = m->iau_inc + (long)n * len, *lo = n > 0 ? hi - len : hi;
    for (long i = 0; i < len; i++) F[i] = (whi * hi[i] + wlo * lo[i]) / m->iau_dt;
    return 1;
}

static double area_sum(const orc_model *m, const double *x) {
    double s = 0.0;
    for (long i = 0; i < m->ng; i++) s += m->areawts[i] * x[i];
    return s;
}

/* ---- dyn_run.f90:307-373 getomega ----------------------------------
 * On ng columns, arrays columns-fastest: ug, vg, divg (nl levels bottom ->
 * top); bk (nl+1), dbk, ck (nl) and pk (nl+1), dpk, alfa, rlnp (nl) top ->
 * bottom; psg, dlnpsdx, dlnpsdy per column -> dlnpsdt (per column), etadot
 * (nl+1 interfaces top -> bottom) and dlnpdtg = omega/p (bottom -> top).
 * cg, cb, db, workb and workc are per-column scratch (reference work arrays). */
void orc_omega_cols(int nl, long ng, const double *bk, const double *dbk, const double *ck,
                    const double *psg, const double *dpk, const double *alfa, const double *rlnp,
                    const double *ug, const double *vg, const double *divg, const double *dlnpsdx,
                    const double *dlnpsdy, double *dlnpsdt, double *etadot, double *dlnpdtg) {
#define LV(X, k) ((X) + (long)((k) - 1) * ng)
#pragma omp parallel
    {
        double *cg = malloc(sizeof(double) * (nl + 1)), *cb = malloc(sizeof(double) * (nl + 1));
        double *db = malloc(sizeof(double) * (nl + 1)), *wb = malloc(sizeof(double) * (nl + 1));
        double *wc = malloc(sizeof(double) * (nl + 1));
#pragma omp for
        for (long i = 0; i < ng; i++) {
            double ps = psg[i];
            /* 1-based level index k on the scratch arrays */
            for (int k = 1; k <= nl; k++)
                cg[k] = LV(ug, nl + 1 - k)[i] * dlnpsdx[i] + LV(vg, nl + 1 - k)[i] * dlnpsdy[i];
            db[1] = LV(divg, nl)[i] * LV(dpk, 1)[i];
            cb[1] = cg[1] * dbk[0];
            for (int k = 1; k <= nl - 1; k++) {
                db[k + 1] = db[k] + LV(divg, nl - k)[i] * LV(dpk, k + 1)[i];
                cb[k + 1] = cb[k] + cg[k + 1] * dbk[k];
            }
            double dpsdt = -db[nl] / ps - cb[nl];
            dlnpsdt[i] = dpsdt;
            LV(etadot, 1)[i] = 0.0;
            LV(etadot, nl + 1)[i] = 0.0;
            for (int k = 1; k <= nl - 1; k++)
                LV(etadot, k + 1)[i] = -ps * (bk[k] * dpsdt + cb[k]) - db[k];
            wb[1] = LV(alfa, 1)[i] * (LV(divg, nl)[i] * LV(dpk, 1)[i] + ps * cb[1] * dbk[0]);
            for (int k = 2; k <= nl; k++)
                wb[k] = LV(rlnp, k)[i] * (db[k - 1] + ps * cb[k - 1]) +
                        LV(alfa, k)[i] * (LV(divg, nl + 1 - k)[i] * LV(dpk, k)[i] + ps * cg[k] * dbk[k - 1]);
            wc[1] = ps * cg[1] * dbk[0];
            for (int k = 2; k <= nl; k++)
                wc[k] = ps * cg[k] * (dbk[k - 1] + ck[k - 1] * LV(rlnp, k)[i] / LV(dpk, k)[i]);
            for (int k = 1; k <= nl; k++)
                LV(dlnpdtg, nl + 1 - k)[i] = (wc[k] - wb[k]) / LV(dpk, k)[i];
        }
        free(cg); free(cb); free(db); free(wb); free(wc);
    }
#undef LV
}

static void getomega(orc_model *m) {
    orc_omega_cols(m->nlev, m->ng, m->bk, m->dbk, m->ck, m->psg, m->dpk, m->alfa, m->rlnp, m->ug, m->vg,
                   m->divg, m->dlnpsdx, m->dlnpsdy, m->dlnpsdt, m->etadot, m->dlnpdtg);
}

/* ---- dyn_run.f90:375-451 getpresgrad -------------------------------
 * On ng columns, arrays columns-fastest: tvg, dtvdx, dtvdy (nl levels bottom
 * -> top); bk (nl+1), dbk (nl) and pk (nl+1), dpk, alfa, rlnp (nl) top ->
 * bottom; psg, dphisdx/y, dlnpsdx/y per column -> prsgx, prsgy (bottom ->
 * top). cofa, cofb, px3u and px3v are per-column scratch (the reference's
 * work arrays, 1-based level index). */
void orc_presgrad_cols(int nl, long ng, double rd, const double *bk, const double *dbk,
                       const double *psg, const double *pk, const double *dpk, const double *alfa,
                       const double *rlnp, const double *tvg, const double *dtvdx, const double *dtvdy,
                       const double *dphisdx, const double *dphisdy, const double *dlnpsdx,
                       const double *dlnpsdy, double *prsgx, double *prsgy) {
#define LV(X, k) ((X) + (long)((k) - 1) * ng)
#pragma omp parallel
    {
        double *cofa = malloc(sizeof(double) * (nl + 1)), *cofb = malloc(sizeof(double) * (nl + 1));
        double *px3u = malloc(sizeof(double) * (nl + 1)), *px3v = malloc(sizeof(double) * (nl + 1));
#pragma omp for
        for (long i = 0; i < ng; i++) {
            double ps = psg[i];
#define P(k) LV(pk, k)[i]
#define TV(k) LV(tvg, k)[i]
            cofb[1] = -(1. / LV(dpk, 1)[i]) * (LV(alfa, 1)[i] * dbk[0]);
            for (int k = 2; k <= nl; k++)
                cofb[k] = -(1. / LV(dpk, k)[i]) * (bk[k - 1] * LV(rlnp, k)[i] + LV(alfa, k)[i] * dbk[k - 1]);
            for (int k = 1; k <= nl; k++) {
                LV(prsgx, nl - k + 1)[i] = cofb[k] * rd * TV(nl + 1 - k) * ps * dlnpsdx[i];
                LV(prsgy, nl - k + 1)[i] = cofb[k] * rd * TV(nl + 1 - k) * ps * dlnpsdy[i];
            }
            for (int k = 1; k <= nl; k++)
                cofa[k] = -(1. / LV(dpk, k)[i]) *
                          (bk[k] * P(k) / P(k + 1) - bk[k - 1] +
                           LV(rlnp, k)[i] * (bk[k - 1] - P(k) * dbk[k - 1] / LV(dpk, k)[i]));
            cofb[nl] = 0.0;
            cofb[nl - 1] = -rd * (bk[nl] / P(nl + 1) - bk[nl - 1] / P(nl)) * TV(1);
            for (int k = 2; k <= nl - 1; k++)
                cofb[nl - k] = cofb[nl + 1 - k] - rd * (bk[nl + 1 - k] / P(nl + 2 - k) - bk[nl - k] / P(nl + 1 - k)) * TV(k);
            px3u[nl] = 0.0;
            px3v[nl] = 0.0;
            px3u[nl - 1] = -rd * LV(rlnp, nl)[i] * LV(dtvdx, 1)[i];
            px3v[nl - 1] = -rd * LV(rlnp, nl)[i] * LV(dtvdy, 1)[i];
            for (int k = 2; k <= nl - 1; k++) {
                px3u[nl - k] = px3u[nl + 1 - k] - rd * LV(rlnp, nl + 1 - k)[i] * LV(dtvdx, k)[i];
                px3v[nl - k] = px3v[nl + 1 - k] - rd * LV(rlnp, nl + 1 - k)[i] * LV(dtvdy, k)[i];
            }
            for (int k = 1; k <= nl; k++) {
                LV(prsgx, nl - k + 1)[i] = LV(prsgx, nl - k + 1)[i] - dphisdx[i] + cofb[k] * ps * dlnpsdx[i] +
                                           px3u[k] - rd * LV(alfa, k)[i] * LV(dtvdx, nl + 1 - k)[i] -
                                           cofa[k] * rd * TV(nl + 1 - k) * ps * dlnpsdx[i];
                LV(prsgy, nl - k + 1)[i] = LV(prsgy, nl - k + 1)[i] - dphisdy[i] + cofb[k] * ps * dlnpsdy[i] +
                                           px3v[k] - rd * LV(alfa, k)[i] * LV(dtvdy, nl + 1 - k)[i] -
                                           cofa[k] * rd * TV(nl + 1 - k) * ps * dlnpsdy[i];
            }
#undef P
#undef TV
        }
        free(cofa); free(cofb); free(px3u); free(px3v);
    }
#undef LV
}

static void getpresgrad(orc_model *m, const double *dtvdx, const double *dtvdy, double *prsgx,
                        double *prsgy) {
    orc_presgrad_cols(m->nlev, m->ng, m->c.rd, m->bk, m->dbk, m->psg, m->pk, m->dpk, m->alfa, m->rlnp,
                      m->tvg, dtvdx, dtvdy, m->dphisdx, m->dphisdy, m->dlnpsdx, m->dlnpsdy, prsgx, prsgy);
}

/* ---- dyn_run.f90:453-476 getvadv -----------------------------------
 * Columns-fastest arrays as orc_vadv_tracers_cols: datag (nl levels bottom ->
 * top), et (nl+1 interfaces top -> bottom), dpk (nl layers top -> bottom) ->
 * vadv (bottom -> top). */
void orc_vadv_cols(int nl, long ng, const double *datag, const double *et, const double *dpk, double *vadv) {
#define LV(X, k) ((X) + (long)((k) - 1) * ng)
#pragma omp parallel for
    for (long i = 0; i < ng; i++) {
        LV(vadv, nl)[i] = (0.5 / LV(dpk, 1)[i]) * LV(et, 2)[i] * (LV(datag, nl - 1)[i] - LV(datag, nl)[i]);
        LV(vadv, 1)[i] = (0.5 / LV(dpk, nl)[i]) * LV(et, nl)[i] * (LV(datag, 1)[i] - LV(datag, 2)[i]);
        for (int k = 2; k <= nl - 1; k++)
            LV(vadv, nl + 1 - k)[i] =
                (0.5 / LV(dpk, k)[i]) * (LV(et, k + 1)[i] * (LV(datag, nl - k)[i] - LV(datag, nl + 1 - k)[i]) +
                                         LV(et, k)[i] * (LV(datag, nl + 1 - k)[i] - LV(datag, nl + 2 - k)[i]));
    }
#undef LV
}

static void getvadv(orc_model *m, const double *datag, double *vadv) {
    orc_vadv_cols(m->nlev, m->ng, datag, m->etadot, m->dpk, vadv);
}

/* ---- dyn_run.f90:478-603 getvadv_tracers (Thuburn 1993 limiter) -----
 * Columns-fastest arrays: datag (nl levels bottom -> top), etadot (nl+1
 * interfaces top -> bottom), dpk (nl layers top -> bottom) -> vadv (bottom -> top). */
void orc_vadv_tracers_cols(int nl, long ng, const double *datag, const double *et, const double *dpk,
                           double *vadv) {
    const double epstiny = 2.2250738585072014e-308; /* tiny(real(8)) */
#define LL(X, k) ((X) + (long)((k) - 1) * ng)
#pragma omp parallel
    {
        double *half = malloc(sizeof(double) * (nl + 1)), *dd = malloc(sizeof(double) * (nl + 1));
#pragma omp for
        for (long i = 0; i < ng; i++) {
#define D(k) LL(datag, k)[i]
            for (int k = 1; k <= nl - 1; k++) {
                half[k] = 0.5 * (D(nl - k) + D(nl + 1 - k));
                dd[k] = D(nl - k) - D(nl + 1 - k);
            }
            half[0] = D(nl);
            half[nl] = D(1);
            if (D(nl) >= 0) dd[0] = D(nl) - fmax(0., 2. * D(nl) - D(nl - 1));
            else dd[0] = D(nl) - fmin(0., 2. * D(nl) - D(nl - 1));
            if (D(1) >= 0) dd[nl] = fmax(0., 2. * D(1) - D(2)) - D(1);
            else dd[nl] = fmin(0., 2. * D(1) - D(2)) - D(1);
            for (int k = 1; k <= nl - 1; k++)
                if (fabs(dd[k]) < epstiny) dd[k] = copysign(epstiny, dd[k]);
            for (int k = 1; k <= nl - 1; k++) {
                double phi;
                if (LL(et, k + 1)[i] > 0.) {
                    phi = dd[k - 1] / dd[k];
                    half[k] = D(nl + 1 - k) + (phi + fabs(phi)) / (1. + fabs(phi)) * (half[k] - D(nl + 1 - k));
                } else {
                    phi = dd[k + 1] / dd[k];
                    half[k] = D(nl - k) + (phi + fabs(phi)) / (1. + fabs(phi)) * (half[k] - D(nl - k));
                }
            }
            for (int k = 1; k <= nl; k++)
                LL(vadv, nl + 1 - k)[i] = (1. / LL(dpk, k)[i]) *
                                          ((half[k] * LL(et, k + 1)[i] - half[k - 1] * LL(et, k)[i]) +
                                           (D(nl + 1 - k) * (LL(et, k)[i] - LL(et, k + 1)[i])));
#undef D
        }
        free(half); free(dd);
    }
#undef LL
}

static void getvadv_tracers(orc_model *m, const double *datag, double *vadv) {
    orc_vadv_tracers_cols(m->nlev, m->ng, datag, m->etadot, m->dpk, vadv);
}

/* ---- dyn_run.f90:33-305 getdyntend --------------------------------- */
void orc_getdyntend(orc_model *m, int rkstage, int inverse_only, ocplx *dvrt, ocplx *ddiv,
                    ocplx *dtv, ocplx *dtrac, ocplx *dlnps) {
    const orc_sht *h = m->sht;
    int nl = m->nlev, nlm = m->nlm, ntrac = m->ntrac;
    long ng = m->ng;
    double rerth = m->c.rerth;
    double *prsgx = m->w[0], *prsgy = m->w[1], *vadvu = m->w[2], *vadvv = m->w[3];
    double *vadvt = m->w[4], *vadvq = m->w[5], *dtvdx = m->w[6], *dtvdy = m->w[7];

    /* dyn_run.f90:101-117 */
    orc_getuv(h, m->vrt, m->div, m->ug, m->vg, nl, rerth);
    orc_spectogrd(h, m->vrt, m->vrtg, nl);
    orc_spectogrd(h, m->div, m->divg, nl);
    orc_spectogrd(h, m->tv, m->tvg, nl);
    orc_getgrad(h, m->tv, dtvdx, dtvdy, nl, rerth);
    if (ntrac > 0) orc_spectogrd(h, m->trac, m->tracg, nl * ntrac);
    /* 114-116: vorticity gradient for vorticity confinement */
    if (m->vc_on && !inverse_only) orc_getgrad(h, m->vrt, m->vcx, m->vcy, nl, rerth);
    /* 126-131 */
    orc_spectogrd(h, m->lnps, m->lnpsg, 1);
    calc_press_cols(nl, ng, m->ak, m->bk, m->kappa, m->lnpsg, m->psg, m->pk, m->dpk, m->alfa,
                    m->rlnp, m->prs);
    /* 135-153: dry-mass diagnostic, pdryini set on first use */
    if (rkstage == 1) {
        double pmoist = 0.0;
        if (ntrac > 0)
            for (int k = 1; k <= nl; k++)
                for (int nt = 0; nt < ntrac; nt++) {
                    const double *q = m->tracg + ((long)nt * nl + (k - 1)) * ng;
                    const double *dp = L(m->dpk, nl - k + 1);
                    for (long i = 0; i < ng; i++) pmoist += m->areawts[i] * dp[i] * q[i];
                }
        double pdry = area_sum(m, m->psg) - pmoist;
        if (m->pdryini == 0.0) m->pdryini = pdry;
    }
    /* 158-166 */
    orc_getgrad(h, m->lnps, m->dlnpsdx, m->dlnpsdy, 1, rerth);
    getomega(m);
    m->have_grid = 1;
    if (inverse_only) return;
    /* 179-188 */
    orc_grdtospec(h, m->dlnpsdt, dlnps, 1);
    getpresgrad(m, dtvdx, dtvdy, prsgx, prsgy);
    getvadv(m, m->ug, vadvu);
    getvadv(m, m->vg, vadvv);
    getvadv(m, m->tvg, vadvt);
    /* 191-214: planetary vorticity and energy conversion */
    double kappa = m->kappa, cvcp = m->c.cvap / m->c.cp - 1.0;
    double twoom = 2. * m->c.omega;
#pragma omp parallel for
    for (long i = 0; i < ng * nl; i++) {
        double t = kappa * m->dlnpdtg[i] * m->tvg[i];
        if (ntrac > 0) t = t / (1. + cvcp * m->tracg[i]);
        vadvq[i] = t;
    }
    /* 220-272 */
#pragma omp parallel for
    for (long idx = 0; idx < ng * nl; idx++) {
        long i = idx % ng;
        double f = twoom * m->sinlat[i];
        vadvv[idx] -= prsgy[idx];
        vadvu[idx] -= prsgx[idx];
        double u = m->ug[idx], v = m->vg[idx], z = m->vrtg[idx];
        m->w[8][idx] = -u * dtvdx[idx] - v * dtvdy[idx] - vadvt[idx] + vadvq[idx];
        prsgx[idx] = u * (z + f) + vadvv[idx];
        prsgy[idx] = v * (z + f) - vadvu[idx];
        m->w[9][idx] = 0.5 * (u * u + v * v);
        if (m->vc_on) {
            /* 230-248: up-gradient vorticity flux vc grad(vrt)/|grad(vrt)| |vrt| with
             * vc = vcamp, times (1 + vfact_svc grd_svc) with the svc pattern */
            double gx = m->vcx[idx], gy = m->vcy[idx], ag = sqrt(gx * gx + gy * gy);
            double vc = m->sc.vcamp;
            if (m->stoch_on[0]) vc = (1. + m->pat[0].vfact[idx / ng] * m->pgrd[0][i]) * vc;
            if (ag > FLT_MIN) {
                prsgx[idx] += vc * (gx / ag) * fabs(z);
                prsgy[idx] += vc * (gy / ag) * fabs(z);
            }
        }
    }
    orc_grdtospec(h, m->w[8], dtv, nl);
    orc_getvrtdivspec(h, prsgx, prsgy, ddiv, dvrt, nl, rerth); /* swapped outputs */
    ocplx *ke = malloc(sizeof(ocplx) * nlm * nl);
    orc_grdtospec(h, m->w[9], ke, nl);
    for (long idx = 0; idx < (long)nlm * nl; idx++) {
        int k = idx % nlm;
        dvrt[idx] = -dvrt[idx];
        ddiv[idx] = ddiv[idx] - (m->lap[k] / (rerth * rerth)) * ke[idx];
    }
    free(ke);
    /* 279-291: tracers */
    for (int nt = 0; nt < ntrac; nt++) {
        double *q = m->tracg + (long)nt * nl * ng;
        getvadv_tracers(m, q, vadvq);
        orc_getgrad(h, m->trac + (long)nt * nl * nlm, dtvdx, dtvdy, nl, rerth);
#pragma omp parallel for
        for (long idx = 0; idx < ng * nl; idx++)
            prsgx[idx] = -m->ug[idx] * dtvdx[idx] - m->vg[idx] * dtvdy[idx] - vadvq[idx];
        orc_grdtospec(h, prsgx, dtrac + (long)nt * nl * nlm, nl);
    }
}

/* ---- phy_run_hs.f90:22-84 ------------------------------------------ */
void orc_hs_phytend(orc_model *m, ocplx *dvrt, ocplx *ddiv, ocplx *dtv) {
    int nl = m->nlev;
    long ng = m->ng;
    const double sigbot = SGL(0.7), delthz = 10., tempstrat = 200.;
    const double kdrag = SGL(1.f / (1.f * 86400.f)), krada = SGL(1.f / (40.f * 86400.f));
    const double kradb = SGL(1.f / (4.f * 86400.f)), p0 = 1.e5, deltmp = 60.;
    const double expo = SGL(2.f / 7.f);
    double *ft = m->w[0], *fv = m->w[1], *fd = m->w[2];
#pragma omp parallel for
    for (long idx = 0; idx < ng * nl; idx++) {
        long i = idx % ng;
        double prs = m->prs[idx], sl = m->sinlat[i], cl = m->coslat[i];
        double blprof = prs / m->psg[i];
        double teq = pow(prs / p0, expo) * (315. - deltmp * sl * sl - delthz * log(prs / p0) * cl * cl);
        blprof = (blprof - sigbot) / (1. - sigbot);
        if (blprof < 0) blprof = 0;
        if (teq < tempstrat) teq = tempstrat;
        ft[idx] = (krada + (kradb - krada) * blprof * cl * cl * cl * cl) * (teq - m->tvg[idx]);
        fv[idx] = -(blprof * kdrag) * m->vrtg[idx];
        fd[idx] = -(blprof * kdrag) * m->divg[idx];
    }
    orc_grdtospec(m->sht, ft, dtv, nl);
    orc_grdtospec(m->sht, fv, dvrt, nl);
    orc_grdtospec(m->sht, fd, ddiv, nl);
}

/* ---- phy_run_simple.f90:28-196 getphytend (simple moist physics) --------
 * On the model's current grid fields (after getdyntend(3, inverse_only)):
 * each column is handed to simple_physics top -> bottom with T = Tv/(1+fv q)
 * (50-63), the updated column is turned back into Tv (92-94), grid tendencies
 * are the changes over the step (95-101), analysed and divided by dt
 * (160-171). Tracer 1 is specific humidity; other tracers get no tendency. */
void orc_simple_phytend(orc_model *m, ocplx *dvrt, ocplx *ddiv, ocplx *dtv, ocplx *dtrac) {
    int nl = m->nlev, nlm = m->nlm;
    long ng = m->ng;
    double dt = m->c.dt;
    const double qmin = SGL(1.e-10), fv = m->c.rv / m->c.rd - 1.0;
    double *dT = m->w[0], *du = m->w[1], *dv = m->w[2], *dq = m->w[3];
    const double *q0 = m->tracg; /* tracer 1, levels bottom -> top */
#pragma omp parallel
    {
        double *gt = malloc(sizeof(double) * nl), *gq = malloc(sizeof(double) * nl);
        double *gu = malloc(sizeof(double) * nl), *gv = malloc(sizeof(double) * nl);
        double *prsl = malloc(sizeof(double) * nl), *prsi = malloc(sizeof(double) * (nl + 1));
        double *delp = malloc(sizeof(double) * nl), *rdelp = malloc(sizeof(double) * nl);
#pragma omp for
        for (long i = 0; i < ng; i++) {
            for (int k = 1; k <= nl; k++) {
                int c = nl - k; /* top -> bottom position of level k */
                double q = L(q0, k)[i];
                gq[c] = q;
                gu[c] = L(m->ug, k)[i];
                gv[c] = L(m->vg, k)[i];
                gt[c] = L(m->tvg, k)[i] / (1. + fv * fmax(qmin, q));
                prsl[c] = L(m->prs, k)[i];
            }
            for (int k = 1; k <= nl + 1; k++) prsi[k - 1] = L(m->pk, k)[i];
            for (int k = 0; k < nl; k++) {
                delp[k] = prsi[k + 1] - prsi[k];
                rdelp[k] = 1. / delp[k];
            }
            double precl;
            orc_simple_column(nl, dt, m->lats[i / m->nlon], gt, gq, gu, gv, prsl, prsi, delp, rdelp,
                              m->psg[i], &precl, m->c.simple_test, m->c.turbflux);
            m->precip[i] = precl * 1000.;
            double pw = 0.0;
            for (int k = 1; k <= nl; k++) {
                int c = nl - k;
                double tv = gt[c] * (1. + fv * fmax(qmin, gq[c]));
                L(dT, k)[i] = tv - L(m->tvg, k)[i];
                L(dq, k)[i] = gq[c] - L(q0, k)[i];
                L(du, k)[i] = gu[c] - L(m->ug, k)[i];
                L(dv, k)[i] = gv[c] - L(m->vg, k)[i];
                pw += L(m->dpk, nl - k + 1)[i] * (L(dq, k)[i] + L(q0, k)[i]);
            }
            m->pwat[i] = pw / m->c.grav;
        }
        free(gt); free(gq); free(gu); free(gv); free(prsl); free(prsi); free(delp); free(rdelp);
    }
    orc_grdtospec(m->sht, dT, dtv, nl);
    orc_getvrtdivspec(m->sht, du, dv, dvrt, ddiv, nl, m->c.rerth);
    memset(dtrac, 0, sizeof(ocplx) * nlm * nl * m->ntrac);
    orc_grdtospec(m->sht, dq, dtrac, nl);
    for (long i = 0; i < (long)nlm * nl; i++) {
        dtv[i] /= dt; dvrt[i] /= dt; ddiv[i] /= dt; dtrac[i] /= dt;
    }
}

/* ---- run.f90:194-565 advance ---------------------------------------- */
static void semi_implicit(orc_model *m, int stage, ocplx **dd, ocplx **dT, ocplx **dp,
                          const ocplx *div0, const ocplx *tv0, const ocplx *lnps0,
                          const double *ca, const double *caa, double cimp);

void orc_advance(orc_model *m, int nsteps) {
    int nl = m->nlev, nlm = m->nlm, ntrac = m->ntrac;
    double dt = m->c.dt;
    size_t s3 = sizeof(ocplx) * nlm * nl, st = s3 * (ntrac > 0 ? ntrac : 1), s1 = sizeof(ocplx) * nlm;
    ocplx *vrt0 = malloc(s3), *div0 = malloc(s3), *tv0 = malloc(s3), *tr0 = malloc(st), *lp0 = malloc(s1);
    ocplx *dv[3], *dd[3], *dT[3], *dq[3], *dp[3];
    for (int s = 0; s < 3; s++) {
        dv[s] = malloc(s3); dd[s] = malloc(s3); dT[s] = malloc(s3); dq[s] = malloc(st); dp[s] = malloc(s1);
    }
    /* explicit / implicit coefficient rows per stage (run.f90:313-489) */
    const double ca[3][3] = {{A21, 0, 0}, {A31, A32, 0}, {B1, B2, B3}};
    const double caa[3][3] = {{AA21, 0, 0}, {AA31, AA32, 0}, {BB1, BB2, BB3}};
    const double cimp[3] = {AA22, AA33, BB4};
    long len = state_len(m);
    ocplx *F = m->iau_n ? malloc(sizeof(ocplx) * len) : NULL;
    for (int step = 0; step < nsteps; step++) {
        /* run.f90:250-258: the IAU forcing of this step, at its start time */
        int iau = F ? iau_forcing(m, m->time, F) : 0;
        /* run.f90:272-292: advance the random patterns used this step */
        if (m->c.tpat_amp > 0) orc_pattern_advance(m);
        stoch_patterns(m);
        memcpy(vrt0, m->vrt, s3); memcpy(div0, m->div, s3); memcpy(tv0, m->tv, s3);
        if (ntrac > 0) memcpy(tr0, m->trac, st);
        memcpy(lp0, m->lnps, s1);
        for (int s = 0; s < 3; s++) {
            orc_getdyntend(m, s + 1, 0, dv[s], dd[s], dT[s], dq[s], dp[s]);
            /* run.f90:302-309,363-370,429-436: the IAU forcing is added to
             * every stage's tendencies (before the linear terms come off) */
            if (iau) {
                long s3n = (long)nlm * nl;
                const ocplx *fv = F, *fd = F + s3n, *ft = F + 2 * s3n, *fq = F + 3 * s3n;
                const ocplx *fp = F + (long)(3 + ntrac) * s3n;
                for (long i = 0; i < s3n; i++) {
                    dv[s][i] += fv[i];
                    dd[s][i] += fd[i];
                    dT[s][i] += ft[i];
                }
                for (long i = 0; i < s3n * ntrac; i++) dq[s][i] += fq[i];
                for (int k = 0; k < nlm; k++) dp[s][k] += fp[k];
            }
            /* vorticity and tracers: always explicit */
            for (long i = 0; i < (long)nlm * nl; i++) {
                ocplx acc = 0.0;
                for (int r = 0; r <= s; r++) acc += ca[s][r] * dv[r][i];
                m->vrt[i] = vrt0[i] + dt * acc;
            }
            for (long i = 0; i < (long)nlm * nl * ntrac; i++) {
                ocplx acc = 0.0;
                for (int r = 0; r <= s; r++) acc += ca[s][r] * dq[r][i];
                m->trac[i] = tr0[i] + dt * acc;
            }
            if (m->c.explicit_) {
                for (long i = 0; i < (long)nlm * nl; i++) {
                    ocplx a1 = 0.0, a2 = 0.0;
                    for (int r = 0; r <= s; r++) { a1 += ca[s][r] * dd[r][i]; a2 += ca[s][r] * dT[r][i]; }
                    m->div[i] = div0[i] + dt * a1;
                    m->tv[i] = tv0[i] + dt * a2;
                }
                for (int k = 0; k < nlm; k++) {
                    ocplx a = 0.0;
                    for (int r = 0; r <= s; r++) a += ca[s][r] * dp[r][k];
                    m->lnps[k] = lp0[k] + dt * a;
                }
            } else {
                semi_implicit(m, s, dd, dT, dp, div0, tv0, lp0, ca[s], caa[s], cimp[s]);
            }
        }
        /* run.f90:493-502 implicit damping */
#pragma omp parallel for
        for (int k = 0; k < nl; k++)
            for (int n = 0; n < nlm; n++) {
                long i = (long)k * nlm + n;
                double fvd = 1. - (m->disspec[n] * m->diff_prof[k] - m->dmp_prof[k]) * dt;
                double ft = 1. - m->disspec[n] * m->diff_prof[k] * dt;
                m->vrt[i] /= fvd;
                m->div[i] /= fvd;
                m->tv[i] /= ft;
                for (int nt = 0; nt < ntrac; nt++) m->trac[(long)nt * nl * nlm + i] /= ft;
            }
        /* run.f90:509-538 time-split physics */
        if (m->c.physics != 0) {
            orc_getdyntend(m, 3, 1, NULL, NULL, NULL, NULL, NULL);
            /* run.f90:517-518 */
            if (m->stoch_on[3]) stoch_forcing(m);
            if (m->c.physics == 1) {
                orc_hs_phytend(m, dv[1], dd[1], dT[1]);
                for (long i = 0; i < (long)nlm * nl; i++) {
                    m->vrt[i] += dt * dv[1][i];
                    m->div[i] += dt * dd[1][i];
                    m->tv[i] += dt * dT[1][i];
                }
                for (int k = 0; k < nlm; k++) dp[1][k] = 0.0;
            } else if (m->c.physics == 2) {
                orc_simple_phytend(m, dv[1], dd[1], dT[1], dq[1]);
                for (long i = 0; i < (long)nlm * nl; i++) {
                    m->vrt[i] += dt * dv[1][i];
                    m->div[i] += dt * dd[1][i];
                    m->tv[i] += dt * dT[1][i];
                }
                for (long i = 0; i < (long)nlm * nl * ntrac; i++) m->trac[i] += dt * dq[1][i];
                for (int k = 0; k < nlm; k++) dp[1][k] = 0.0;
            }
            if (m->c.massfix) {
                /* dyn_run.f90:605-634; pwat is 0 for Held-Suarez (phy_data_stub.f90:17)
                 * and the post-physics column water for the simple physics */
                double pwatg = m->c.physics == 2 ? area_sum(m, m->pwat) : 0.0;
                double pmean = area_sum(m, m->psg);
                double pcorr = (m->pdryini + m->c.grav * pwatg) / pmean;
                double *g = m->w[3];
                for (long i = 0; i < m->ng; i++) g[i] = log(m->psg[i] * pcorr);
                ocplx *ws = malloc(s1);
                orc_grdtospec(m->sht, g, ws, 1);
                for (int k = 0; k < nlm; k++) dp[1][k] = (ws[k] - m->lnps[k]) / dt;
                free(ws);
            }
            for (int k = 0; k < nlm; k++) m->lnps[k] += dt * dp[1][k];
            if (m->stoch_on[3]) stoch_apply(m);
        }
        /* stochastic temperature forcing pattern/tau, in the place of the
         * additive-noise tvforcing of run.f90:539-551 */
        if (m->c.tpat_amp > 0)
            for (int k = 0; k < nl; k++)
                for (int n = 0; n < nlm; n++)
                    m->tv[(long)k * nlm + n] += dt * m->tpat_vfact[k] * m->tpat[n] / m->c.tpat_tau;
        m->time += dt; /* run.f90:153 */
    }
    free(F);
    free(vrt0); free(div0); free(tv0); free(tr0); free(lp0);
    for (int s = 0; s < 3; s++) {
        free(dv[s]); free(dd[s]); free(dT[s]); free(dq[s]); free(dp[s]);
    }
}

/* One semi-implicit stage of run.f90:326-350 / 386-416 / 450-488 for every
 * spectral coefficient: remove the linear terms from the stage tendency,
 * form the explicit update with the (a, aa) row of the tableau, solve for
 * the new divergence with d_hyb_m(:,:,n+1,stage) and back-substitute. The
 * linear terms of each stage are kept in the model for the later stages. */
static void semi_implicit(orc_model *m, int stage, ocplx **dd, ocplx **dT, ocplx **dp,
                          const ocplx *div0, const ocplx *tv0, const ocplx *lnps0,
                          const double *ca, const double *caa, double cimp) {
    int nl = m->nlev, nlm = m->nlm;
    double dt = m->c.dt;
    ocplx **g_lin_d = m->lin_d, **g_lin_t = m->lin_t, **g_lin_p = m->lin_p;
    const double *A = m->amhyb, *Bm = m->bmhyb, *sv = m->svhyb, *tor = m->tor_hyb;
#pragma omp parallel
    {
        ocplx *tmp = malloc(sizeof(ocplx) * nl), *rhs = malloc(sizeof(ocplx) * nl);
        ocplx *tvc = malloc(sizeof(ocplx) * nl), *dvc = malloc(sizeof(ocplx) * nl);
#pragma omp for schedule(static)
        for (int n = 0; n < nlm; n++) {
            double lapn = m->lap[n];
            /* linear terms of the stage's input state */
            const ocplx *tvin = (stage == 0) ? tv0 : m->tv;
            const ocplx *dvin = (stage == 0) ? div0 : m->div;
            ocplx lpin = (stage == 0) ? lnps0[n] : m->lnps[n];
            for (int k = 0; k < nl; k++) { tvc[k] = tvin[(long)k * nlm + n]; dvc[k] = dvin[(long)k * nlm + n]; }
            ocplx dpl = 0.0;
            for (int i = 0; i < nl; i++) {
                ocplx a = 0.0, b = 0.0;
                for (int j = 0; j < nl; j++) { a += A[i + j * nl] * tvc[j]; b += Bm[i + j * nl] * dvc[j]; }
                ocplx ddl = -lapn * (a + tor[i] * lpin);
                ocplx dtl = -b;
                g_lin_d[stage][(long)i * nlm + n] = ddl;
                g_lin_t[stage][(long)i * nlm + n] = dtl;
                dd[stage][(long)i * nlm + n] -= ddl;
                dT[stage][(long)i * nlm + n] -= dtl;
                dpl -= sv[i] * dvc[i];
            }
            g_lin_p[stage][n] = dpl;
            dp[stage][n] -= dpl;
            /* explicit part */
            for (int i = 0; i < nl; i++) {
                long id = (long)i * nlm + n;
                ocplx a1 = 0.0, a2 = 0.0;
                for (int r = 0; r <= stage; r++) {
                    a1 += ca[r] * dd[r][id] + caa[r] * g_lin_d[r][id];
                    a2 += ca[r] * dT[r][id] + caa[r] * g_lin_t[r][id];
                }
                dvc[i] = div0[id] + dt * a1;
                tvc[i] = tv0[id] + dt * a2;
            }
            ocplx a3 = 0.0;
            for (int r = 0; r <= stage; r++) a3 += ca[r] * dp[r][n] + caa[r] * g_lin_p[r][n];
            ocplx lpn = lnps0[n] + dt * a3;
            if (stage < 2 || fabs(cimp) > 1.e-5) {
                /* implicit solve */
                for (int i = 0; i < nl; i++) {
                    ocplx a = 0.0;
                    for (int j = 0; j < nl; j++) a += A[i + j * nl] * tvc[j];
                    rhs[i] = dvc[i] - cimp * lapn * dt * (a + tor[i] * lpn);
                }
                const double *D = m->d_hyb_m + ((long)stage * (m->T + 1) + m->degree[n]) * nl * nl;
                for (int i = 0; i < nl; i++) {
                    ocplx a = 0.0;
                    for (int j = 0; j < nl; j++) a += D[i + j * nl] * rhs[j];
                    tmp[i] = a;
                }
                for (int i = 0; i < nl; i++) {
                    ocplx b = 0.0;
                    for (int j = 0; j < nl; j++) b += Bm[i + j * nl] * tmp[j];
                    tvc[i] -= cimp * dt * b;
                    lpn -= cimp * dt * sv[i] * tmp[i];
                }
                for (int i = 0; i < nl; i++) dvc[i] = tmp[i];
            }
            for (int i = 0; i < nl; i++) {
                m->div[(long)i * nlm + n] = dvc[i];
                m->tv[(long)i * nlm + n] = tvc[i];
            }
            m->lnps[n] = lpn;
        }
        free(tmp); free(rhs); free(tvc); free(dvc);
    }
}

/* ---- accessors ------------------------------------------------------ */
void orc_get_grid(const orc_model *m, int which, double *out) {
    long ng = m->ng, n3 = ng * m->nlev;
    const double *src = NULL;
    long n = n3;
    switch (which) {
    case 0: src = m->ug; break;
    case 1: src = m->vg; break;
    case 2: src = m->vrtg; break;
    case 3: src = m->divg; break;
    case 4: src = m->tvg; break;
    case 5: src = m->dlnpdtg; break;
    case 6: src = m->etadot; n = n3 + ng; break;
    case 7: src = m->psg; n = ng; break;
    case 8: src = m->dlnpsdt; n = ng; break;
    case 9: src = m->prs; break;
    case 10: src = m->pk; n = n3 + ng; break;
    case 11: src = m->dpk; break;
    case 12: src = m->alfa; break;
    case 13: src = m->rlnp; break;
    case 14: src = m->dlnpsdx; n = ng; break;
    case 15: src = m->dlnpsdy; n = ng; break;
    case 16: src = m->tracg; n = n3 * (m->ntrac > 0 ? m->ntrac : 0); break;
    case 17: src = m->precip; n = ng; break;
    case 18: src = m->pwat; n = ng; break;
    default: return;
    }
    if (n > 0) memcpy(out, src, sizeof(double) * n);
}

void orc_get_semimp(const orc_model *m, double *amhyb, double *bmhyb, double *svhyb,
                    double *tor_hyb, double *d_hyb_m) {
    int nl = m->nlev;
    if (amhyb) memcpy(amhyb, m->amhyb, sizeof(double) * nl * nl);
    if (bmhyb) memcpy(bmhyb, m->bmhyb, sizeof(double) * nl * nl);
    if (svhyb) memcpy(svhyb, m->svhyb, sizeof(double) * nl);
    if (tor_hyb) memcpy(tor_hyb, m->tor_hyb, sizeof(double) * nl);
    if (d_hyb_m) memcpy(d_hyb_m, m->d_hyb_m, sizeof(double) * nl * nl * (m->T + 1) * 3);
}

void orc_get_damp(const orc_model *m, double *disspec, double *diff_prof, double *dmp_prof) {
    if (disspec) memcpy(disspec, m->disspec, sizeof(double) * m->nlm);
    if (diff_prof) memcpy(diff_prof, m->diff_prof, sizeof(double) * m->nlev);
    if (dmp_prof) memcpy(dmp_prof, m->dmp_prof, sizeof(double) * m->nlev);
}

void orc_step_diag(const orc_model *m, double *out4) {
    long ng = m->ng, n3 = ng * m->nlev;
    double spdmax = 0.0, pmin = 1e300, pmax = -1e300, pst = 0.0;
    for (long i = 0; i < n3; i++) {
        double s = sqrt(m->ug[i] * m->ug[i] + m->vg[i] * m->vg[i]);
        if (s > spdmax) spdmax = s;
    }
    for (long i = 0; i < ng; i++) {
        double p = m->psg[i];
        if (p < pmin) pmin = p;
        if (p > pmax) pmax = p;
        double t = 36. * p * m->dlnpsdt[i];
        pst += t * t * m->areawts[i];
    }
    out4[0] = spdmax; out4[1] = pmin; out4[2] = pmax; out4[3] = sqrt(pst);
}
