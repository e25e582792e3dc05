/*
 * FP64 CPU restatement of the spherical-harmonic transforms behind
 * src/shtns.f90 (TEST INFRASTRUCTURE ONLY -- see oracle.h).
 *
 * What is restated and from where:
 *  - grid metadata of shtns_init (shtns.f90:90-156): Gauss nodes north->south,
 *    lats = asin(cos colat), gauwts symmetric, areawts summing to 1,
 *    degree/order m-major, lap = -n(n+1), invlap(1) = 0.
 *  - SHTns (external, version unpinned, not in the image) with
 *    SHT_ORTHONORMAL + SHT_NO_CS_PHASE on a Gauss grid (shtns.f90:110-111):
 *      f = sum_n q_n0 Y_n0 + 2 Re sum_{m>0} q_nm Y_nm,
 *      Y_nm = Pbar_n^m(cos th) e^{i m lambda}, int |Y|^2 dOmega = 1.
 *    sh_to_spat / spat_to_sh are exact Gauss quadrature + real DFT.
 *    sphtor_to_spat(S,T) -> (Vt,Vp) with Vt = dS/dth + (1/sin th) dT/dphi,
 *    Vp = (1/sin th) dS/dphi - dT/dth, and spat_to_sphtor its quadrature
 *    inverse. These are evaluated here through U = u cos(lat), V = v cos(lat)
 *    and the recurrence (1-mu^2) dP_n/dmu = -n eps_{n+1} P_{n+1} + (n+1) eps_n P_{n-1},
 *    which gives the same quadrature sums node by node (multiplying the
 *    colatitude derivatives by sin th = cos(lat) turns dP/dth into these
 *    neighbouring-degree terms).
 *  - getuv / getvrtdivspec / getgrad scalings and role swaps
 *    (shtns.f90:259-261, 310, 339-351).
 *
 * Pinned by: scipy roots_legendre (nodes/weights), scipy/mpmath Pbar_n^m
 * fixtures, orthonormality, brute-force DFT, solid-body rotation, analytic
 * gradients of Y_nm, and an independent 4-contraction vector quadrature in the
 * tests. SHTns's polar optimisation (threshold 1e-10, shtns.f90:76) is not
 * restated: values it would skip are < 1e-10 of the column maximum.
 */
#include "oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

struct orc_sht {
    int nlon, nlat, nhalf, T, nlm, nlme;
    double *mu, *w, *coslat;   /* nlat */
    double *plm;               /* [m][j<nhalf][n-m], n = m..T+1 */
    long *poff;                /* offset of order m in plm */
    double *eps;               /* extended index: eps_n^m, n = m..T+2 stored at ke(m,n) */
    int *eoff;                 /* extended offsets m*(T+3) - m(m-1)/2 for eps (n up to T+2) */
    ocplx *tw;                 /* twiddles exp(-2 pi i t / nlon) */
};

/* ------------------------------------------------------------------ */
/* Gauss-Legendre nodes by Newton iteration on P_nlat(cos theta) in the
 * colatitude theta (FP64); sin(theta) is then exact to rounding, which keeps
 * the polar weights w = 2 sin^2 / (n (P_{n-1} - x P_n))^2 accurate. */
void orc_gauss(int nlat, double *mu, double *w) {
    int n = nlat;
    for (int i = 0; i < (nlat + 1) / 2; i++) {
        double th = M_PI * (i + 0.75) / (n + 0.5);
        double x = cos(th), p0 = 1.0, p1 = x;
        for (int it = 0; it < 100; it++) {
            x = cos(th);
            p0 = 1.0;
            p1 = x;
            for (int k = 2; k <= n; k++) {
                double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            /* dP_n(cos th)/dth = -n (P_{n-1} - x P_n) / sin th */
            double st = sin(th);
            double dth = -p1 * st / (n * (p0 - x * p1));
            th -= dth;
            if (fabs(dth) < 1e-16) break;
        }
        x = cos(th);
        p0 = 1.0;
        p1 = x;
        for (int k = 2; k <= n; k++) {
            double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
            p0 = p1;
            p1 = p2;
        }
        double st = sin(th), d = n * (p0 - x * p1);
        mu[i] = x;
        mu[nlat - 1 - i] = -x;
        w[i] = w[nlat - 1 - i] = 2.0 * st * st / (d * d);
    }
}

static inline double eps_nm(int n, int m) {
    if (n <= m) return 0.0;
    return sqrt(((double)n * n - (double)m * m) / (4.0 * n * n - 1.0));
}

/* Pbar_n^m(mu), n = m..nmax, orthonormal on the sphere, no CS phase.
 * Extended-exponent recurrence so that high-m values near the poles do not
 * underflow before they can grow; values below 2^-1000 are returned as 0. */
static void legendre_column(int m, int nmax, double mu, double *out) {
    const double BIG = 0x1p500, SMALL = 0x1p-500;
    double s = sqrt((1.0 - mu) * (1.0 + mu));
    double p = sqrt(1.0 / (4.0 * M_PI));
    int e = 0; /* true value = p * BIG^e */
    for (int k = 1; k <= m; k++) {
        p *= sqrt((2.0 * k + 1.0) / (2.0 * k)) * s;
        if (p != 0.0 && fabs(p) < SMALL) { p *= BIG; e--; }
    }
    double pm2 = 0.0, pm1 = p;
    for (int n = m; n <= nmax; n++) {
        double val;
        if (n == m) {
            val = pm1;
        } else {
            val = (mu * pm1 - eps_nm(n - 1, m) * pm2) / eps_nm(n, m);
            pm2 = pm1;
            pm1 = val;
            if (e < 0 && fabs(val) > BIG) { pm1 *= SMALL; pm2 *= SMALL; val *= SMALL; e++; }
        }
        out[n - m] = (e == 0) ? val : (e == -1 ? val * SMALL : 0.0);
    }
}

orc_sht *orc_sht_create(int nlon, int nlat, int ntrunc) {
    orc_sht *h = (orc_sht *)calloc(1, sizeof(orc_sht));
    int T = ntrunc;
    h->nlon = nlon; h->nlat = nlat; h->nhalf = nlat / 2; h->T = T;
    h->nlm = (T + 1) * (T + 2) / 2;
    h->nlme = h->nlm + T + 1;
    h->mu = malloc(sizeof(double) * nlat);
    h->w = malloc(sizeof(double) * nlat);
    h->coslat = malloc(sizeof(double) * nlat);
    orc_gauss(nlat, h->mu, h->w);
    for (int j = 0; j < nlat; j++) h->coslat[j] = sqrt((1.0 - h->mu[j]) * (1.0 + h->mu[j]));
    h->poff = malloc(sizeof(long) * (T + 2));
    long off = 0;
    for (int m = 0; m <= T; m++) { h->poff[m] = off; off += (long)h->nhalf * (T + 2 - m); }
    h->poff[T + 1] = off;
    h->plm = malloc(sizeof(double) * (off > 0 ? off : 1));
#pragma omp parallel for schedule(dynamic)
    for (int m = 0; m <= T; m++)
        for (int j = 0; j < h->nhalf; j++)
            legendre_column(m, T + 1, h->mu[j], h->plm + h->poff[m] + (long)j * (T + 2 - m));
    h->eoff = malloc(sizeof(int) * (T + 2));
    int eo = 0;
    for (int m = 0; m <= T; m++) { h->eoff[m] = eo; eo += T + 3 - m; }
    h->eoff[T + 1] = eo;
    h->eps = malloc(sizeof(double) * eo);
    for (int m = 0; m <= T; m++)
        for (int n = m; n <= T + 2; n++) h->eps[h->eoff[m] + n - m] = eps_nm(n, m);
    h->tw = malloc(sizeof(ocplx) * nlon);
    for (int t = 0; t < nlon; t++) h->tw[t] = cexp(-2.0 * M_PI * I * (double)t / nlon);
    return h;
}

void orc_sht_destroy(orc_sht *h) {
    if (!h) return;
    free(h->mu); free(h->w); free(h->coslat); free(h->plm); free(h->poff);
    free(h->eps); free(h->eoff); free(h->tw); free(h);
}

int orc_sht_nlm(const orc_sht *h) { return h->nlm; }

void orc_sht_griddata(const orc_sht *h, double *lats, double *gauwts, double *areawts,
                      int *degree, int *order, double *lap, double *invlap) {
    double tot = 0.0;
    for (int j = 0; j < h->nlat; j++) tot += h->w[j] * h->nlon;
    for (int j = 0; j < h->nlat; j++) {
        if (lats) lats[j] = asin(h->mu[j]);
        if (gauwts) gauwts[j] = h->w[j];
        if (areawts)
            for (int i = 0; i < h->nlon; i++) areawts[(long)j * h->nlon + i] = h->w[j] / tot;
    }
    int k = 0;
    for (int m = 0; m <= h->T; m++)
        for (int n = m; n <= h->T; n++, k++) {
            if (degree) degree[k] = n;
            if (order) order[k] = m;
            if (lap) lap[k] = -(double)n * (n + 1.0);
            if (invlap) invlap[k] = (k == 0) ? 0.0 : -1.0 / ((double)n * (n + 1.0));
        }
}

void orc_sht_legendre(const orc_sht *h, int m, int j, double *out) {
    int K = h->T + 2 - m;
    int jj = j < h->nhalf ? j : h->nlat - 1 - j;
    const double *p = h->plm + h->poff[m] + (long)jj * K;
    for (int i = 0; i < K; i++) {
        double sgn = (j >= h->nhalf && (i & 1)) ? -1.0 : 1.0;
        out[i] = sgn * p[i];
    }
}

/* ------------------------------------------------------------------ */
/* mixed-radix complex DFT, recursive decimation in time; tw holds
 * exp(sign*2 pi i t / N) for the top-level N, sub-levels use strides. */
static int smallest_factor(int n) {
    for (int p = 2; p * p <= n; p++) if (n % p == 0) return p;
    return n;
}

static void dft_rec(int n, const ocplx *in, int stride, ocplx *out, const ocplx *tw, int N,
                    ocplx *tmp) {
    if (n == 1) { out[0] = in[0]; return; }
    int p = smallest_factor(n), m = n / p;
    for (int q = 0; q < p; q++) dft_rec(m, in + (long)q * stride, stride * p, out + q * m, tw, N, tmp);
    int tstep = N / n;
    for (int k = 0; k < m; k++) {
        for (int s = 0; s < p; s++) tmp[s] = out[s * m + k] * tw[((long)s * k * tstep) % N];
        for (int r = 0; r < p; r++) {
            ocplx acc = 0.0;
            for (int s = 0; s < p; s++) acc += tmp[s] * tw[((long)s * r * (N / p)) % N];
            out[r * m + k] = acc;
        }
    }
}

/* forward (sign -1) DFT via the table, backward by conjugation */
static void cdft(int n, ocplx *x, const ocplx *tw, int sign) {
    ocplx *buf = malloc(sizeof(ocplx) * n);
    ocplx *tmp = malloc(sizeof(ocplx) * (n + 1));
    if (sign > 0) for (int i = 0; i < n; i++) x[i] = conj(x[i]);
    dft_rec(n, x, 1, buf, tw, n, tmp);
    if (sign > 0) for (int i = 0; i < n; i++) buf[i] = conj(buf[i]);
    memcpy(x, buf, sizeof(ocplx) * n);
    free(buf); free(tmp);
}

static ocplx *make_tw(int nlon) {
    ocplx *tw = malloc(sizeof(ocplx) * nlon);
    for (int t = 0; t < nlon; t++) tw[t] = cexp(-2.0 * M_PI * I * (double)t / nlon);
    return tw;
}

/* two real rows -> Fourier coefficients m = 0..mmax, F_m = (1/N) sum f e^{-i m lambda} */
static void fft_anal2(int nlon, int mmax, const double *x, const double *y, ocplx *X, ocplx *Y,
                      const ocplx *tw) {
    ocplx *z = malloc(sizeof(ocplx) * nlon);
    for (int i = 0; i < nlon; i++) z[i] = x[i] + I * (y ? y[i] : 0.0);
    cdft(nlon, z, tw, -1);
    double inv = 1.0 / nlon;
    for (int k = 0; k <= mmax; k++) {
        ocplx zk = z[k], zn = conj(z[(nlon - k) % nlon]);
        X[k] = 0.5 * (zk + zn) * inv;
        if (Y) Y[k] = -0.5 * I * (zk - zn) * inv;
    }
    free(z);
}

/* Fourier coefficients m = 0..mmax (mmax < N/2) of two real rows -> rows */
static void fft_synth2(int nlon, int mmax, const ocplx *X, const ocplx *Y, double *x, double *y,
                       const ocplx *tw) {
    ocplx *z = calloc(nlon, sizeof(ocplx));
    z[0] = creal(X[0]) + I * (Y ? creal(Y[0]) : 0.0);
    for (int k = 1; k <= mmax; k++) {
        ocplx xk = X[k], yk = Y ? Y[k] : 0.0;
        z[k] = xk + I * yk;
        z[nlon - k] = conj(xk) + I * conj(yk);
    }
    cdft(nlon, z, tw, +1);
    for (int i = 0; i < nlon; i++) {
        x[i] = creal(z[i]);
        if (y) y[i] = cimag(z[i]);
    }
    free(z);
}

void orc_fft_analysis(int nlon, int mmax, const double *row, ocplx *coef) {
    ocplx *tw = make_tw(nlon);
    fft_anal2(nlon, mmax, row, NULL, coef, NULL, tw);
    free(tw);
}

void orc_fft_synthesis(int nlon, int mmax, const ocplx *coef, double *row) {
    ocplx *tw = make_tw(nlon);
    fft_synth2(nlon, mmax, coef, NULL, row, NULL, tw);
    free(tw);
}

/* ------------------------------------------------------------------ */
/* Legendre stages on one field. ce: extended spectral (n up to T+1),
 * four: [nlat][T+1] Fourier coefficients. */
static void leg_synth(const orc_sht *h, const ocplx *ce, ocplx *four) {
    int T = h->T, nh = h->nhalf, nlat = h->nlat;
    int ke = 0;
    for (int m = 0; m <= T; m++) {
        int K = T + 2 - m;
        for (int j = 0; j < nh; j++) {
            const double *p = h->plm + h->poff[m] + (long)j * K;
            ocplx E = 0.0, O = 0.0;
            for (int i = 0; i < K; i += 2) E += p[i] * ce[ke + i];
            for (int i = 1; i < K; i += 2) O += p[i] * ce[ke + i];
            four[(long)j * (T + 1) + m] = E + O;
            four[(long)(nlat - 1 - j) * (T + 1) + m] = E - O;
        }
        ke += K;
    }
}

/* four already weighted; ce[ke(m,n)] = sum_j P_n^m(mu_j) four[j][m], n up to T+1 */
static void leg_anal(const orc_sht *h, const ocplx *four, ocplx *ce) {
    int T = h->T, nh = h->nhalf, nlat = h->nlat;
    int ke = 0;
    for (int m = 0; m <= T; m++) {
        int K = T + 2 - m;
        for (int i = 0; i < K; i++) ce[ke + i] = 0.0;
        for (int j = 0; j < nh; j++) {
            const double *p = h->plm + h->poff[m] + (long)j * K;
            ocplx fn = four[(long)j * (T + 1) + m], fs = four[(long)(nlat - 1 - j) * (T + 1) + m];
            ocplx S = fn + fs, A = fn - fs;
            for (int i = 0; i < K; i += 2) ce[ke + i] += p[i] * S;
            for (int i = 1; i < K; i += 2) ce[ke + i] += p[i] * A;
        }
        ke += K;
    }
}

/* grid field -> weighted Fourier coefficients; scale_mode 1 divides by cos(lat) */
static void grid_to_four(const orc_sht *h, const double *g, ocplx *four, int scale_mode) {
    int nlat = h->nlat, nlon = h->nlon, T = h->T;
    for (int j = 0; j < nlat; j += 2) {
        int j2 = j + 1 < nlat ? j + 1 : -1;
        fft_anal2(nlon, T, g + (long)j * nlon, j2 >= 0 ? g + (long)j2 * nlon : NULL,
                  four + (long)j * (T + 1), j2 >= 0 ? four + (long)j2 * (T + 1) : NULL, h->tw);
    }
    for (int j = 0; j < nlat; j++) {
        double s = 2.0 * M_PI * h->w[j];
        if (scale_mode) s /= h->coslat[j];
        for (int m = 0; m <= T; m++) four[(long)j * (T + 1) + m] *= s;
    }
}

static void four_to_grid(const orc_sht *h, ocplx *four, double *g, int scale_mode) {
    int nlat = h->nlat, nlon = h->nlon, T = h->T;
    if (scale_mode)
        for (int j = 0; j < nlat; j++) {
            double s = 1.0 / h->coslat[j];
            for (int m = 0; m <= T; m++) four[(long)j * (T + 1) + m] *= s;
        }
    for (int j = 0; j < nlat; j += 2) {
        int j2 = j + 1 < nlat ? j + 1 : -1;
        fft_synth2(nlon, T, four + (long)j * (T + 1), j2 >= 0 ? four + (long)j2 * (T + 1) : NULL,
                   g + (long)j * nlon, j2 >= 0 ? g + (long)j2 * nlon : NULL, h->tw);
    }
}

static inline int kidx(int T, int m, int n) { return m * (2 * T + 3 - m) / 2 + (n - m); }

/* spectral (n<=T) -> extended (n<=T+1, zero at T+1) */
static void spec_to_ext(const orc_sht *h, const ocplx *s, ocplx *ce) {
    int T = h->T, ke = 0;
    for (int m = 0; m <= T; m++) {
        for (int n = m; n <= T; n++) ce[ke + n - m] = s[kidx(T, m, n)];
        ce[ke + T + 1 - m] = 0.0;
        ke += T + 2 - m;
    }
}

static void ext_to_spec(const orc_sht *h, const ocplx *ce, ocplx *s) {
    int T = h->T, ke = 0;
    for (int m = 0; m <= T; m++) {
        for (int n = m; n <= T; n++) s[kidx(T, m, n)] = ce[ke + n - m];
        ke += T + 2 - m;
    }
}

/* (S, Tt) (n <= T) -> extended U-hat, V-hat (n <= T+1) for U = u cos(lat), V = v cos(lat)
 * with u = Vt, v = Vp of sphtor_to_spat(S, Tt). */
static void sphtor_to_uv_ext(const orc_sht *h, const ocplx *S, const ocplx *Tt, ocplx *Ue,
                             ocplx *Ve) {
    int T = h->T, ke = 0;
    for (int m = 0; m <= T; m++) {
        const double *ep = h->eps + h->eoff[m];
        for (int n = m; n <= T + 1; n++) {
            ocplx Sm1 = (n - 1 >= m) ? S[kidx(T, m, n - 1)] : 0.0;
            ocplx Sp1 = (n + 1 <= T) ? S[kidx(T, m, n + 1)] : 0.0;
            ocplx Tm1 = (n - 1 >= m) ? Tt[kidx(T, m, n - 1)] : 0.0;
            ocplx Tp1 = (n + 1 <= T) ? Tt[kidx(T, m, n + 1)] : 0.0;
            ocplx S0 = (n <= T) ? S[kidx(T, m, n)] : 0.0;
            ocplx T0 = (n <= T) ? Tt[kidx(T, m, n)] : 0.0;
            double en = ep[n - m], en1 = ep[n + 1 - m];
            Ue[ke + n - m] = (n - 1.0) * en * Sm1 - (n + 2.0) * en1 * Sp1 + I * (double)m * T0;
            Ve[ke + n - m] = I * (double)m * S0 - (n - 1.0) * en * Tm1 + (n + 2.0) * en1 * Tp1;
        }
        ke += T + 2 - m;
    }
}

/* extended U~, V~ (quadrature of u/cos, v/cos against P_n, n <= T+1) -> vrt, div (n <= T) */
static void uv_ext_to_vrtdiv(const orc_sht *h, const ocplx *Ue, const ocplx *Ve, ocplx *vrt,
                             ocplx *div, double rsphere) {
    int T = h->T, ke = 0;
    for (int m = 0; m <= T; m++) {
        const double *ep = h->eps + h->eoff[m];
        for (int n = m; n <= T; n++) {
            double en = ep[n - m], en1 = ep[n + 1 - m];
            ocplx Um1 = (n - 1 >= m) ? Ue[ke + n - 1 - m] : 0.0;
            ocplx Vm1 = (n - 1 >= m) ? Ve[ke + n - 1 - m] : 0.0;
            ocplx Up1 = Ue[ke + n + 1 - m], Vp1 = Ve[ke + n + 1 - m];
            ocplx U0 = Ue[ke + n - m], V0 = Ve[ke + n - m];
            vrt[kidx(T, m, n)] = (I * (double)m * V0 - n * en1 * Up1 + (n + 1.0) * en * Um1) / rsphere;
            div[kidx(T, m, n)] = (I * (double)m * U0 + n * en1 * Vp1 - (n + 1.0) * en * Vm1) / rsphere;
        }
        ke += T + 2 - m;
    }
}

/* ------------------------------------------------------------------ */
void orc_spectogrd(const orc_sht *h, const ocplx *spec, double *grid, int nfield) {
    long ng = (long)h->nlon * h->nlat;
#pragma omp parallel
    {
        ocplx *ce = malloc(sizeof(ocplx) * h->nlme);
        ocplx *four = malloc(sizeof(ocplx) * h->nlat * (h->T + 1));
#pragma omp for schedule(dynamic)
        for (int f = 0; f < nfield; f++) {
            spec_to_ext(h, spec + (long)f * h->nlm, ce);
            leg_synth(h, ce, four);
            four_to_grid(h, four, grid + f * ng, 0);
        }
        free(ce); free(four);
    }
}

void orc_grdtospec(const orc_sht *h, const double *grid, ocplx *spec, int nfield) {
    long ng = (long)h->nlon * h->nlat;
#pragma omp parallel
    {
        ocplx *ce = malloc(sizeof(ocplx) * h->nlme);
        ocplx *four = malloc(sizeof(ocplx) * h->nlat * (h->T + 1));
#pragma omp for schedule(dynamic)
        for (int f = 0; f < nfield; f++) {
            grid_to_four(h, grid + f * ng, four, 0);
            leg_anal(h, four, ce);
            ext_to_spec(h, ce, spec + (long)f * h->nlm);
        }
        free(ce); free(four);
    }
}

/* sphtor_to_spat(S,Tt) -> u = Vt, v = Vp, then scaled by 'scale' */
static void sphtor_synth(const orc_sht *h, const ocplx *S, const ocplx *Tt, double *u, double *v,
                         double scale, ocplx *ce, ocplx *ce2, ocplx *four) {
    long ng = (long)h->nlon * h->nlat;
    sphtor_to_uv_ext(h, S, Tt, ce, ce2);
    leg_synth(h, ce, four);
    four_to_grid(h, four, u, 1);
    leg_synth(h, ce2, four);
    four_to_grid(h, four, v, 1);
    if (scale != 1.0)
        for (long i = 0; i < ng; i++) { u[i] *= scale; v[i] *= scale; }
}

void orc_getuv(const orc_sht *h, const ocplx *vrt, const ocplx *div, double *u, double *v,
               int nfield, double rsphere) {
    long ng = (long)h->nlon * h->nlat;
    int nlm = h->nlm;
#pragma omp parallel
    {
        ocplx *S = malloc(sizeof(ocplx) * nlm), *Tt = malloc(sizeof(ocplx) * nlm);
        ocplx *ce = malloc(sizeof(ocplx) * h->nlme), *ce2 = malloc(sizeof(ocplx) * h->nlme);
        ocplx *four = malloc(sizeof(ocplx) * h->nlat * (h->T + 1));
#pragma omp for schedule(dynamic)
        for (int f = 0; f < nfield; f++) {
            /* shtns.f90:259-261: S = invlap*rsphere*vrt, T = invlap*rsphere*div */
            int k = 0;
            for (int m = 0; m <= h->T; m++)
                for (int n = m; n <= h->T; n++, k++) {
                    double il = (k == 0) ? 0.0 : -1.0 / ((double)n * (n + 1.0));
                    S[k] = il * rsphere * vrt[(long)f * nlm + k];
                    Tt[k] = il * rsphere * div[(long)f * nlm + k];
                }
            sphtor_synth(h, S, Tt, u + f * ng, v + f * ng, 1.0, ce, ce2, four);
        }
        free(S); free(Tt); free(ce); free(ce2); free(four);
    }
}

void orc_getgrad(const orc_sht *h, const ocplx *spec, double *gx, double *gy, int nfield,
                 double rsphere) {
    long ng = (long)h->nlon * h->nlat;
    int nlm = h->nlm;
#pragma omp parallel
    {
        ocplx *S = calloc(nlm, sizeof(ocplx));
        ocplx *ce = malloc(sizeof(ocplx) * h->nlme), *ce2 = malloc(sizeof(ocplx) * h->nlme);
        ocplx *four = malloc(sizeof(ocplx) * h->nlat * (h->T + 1));
#pragma omp for schedule(dynamic)
        for (int f = 0; f < nfield; f++)
            /* shtns.f90:339-351: sphtor_to_spat(0, f) / rsphere */
            sphtor_synth(h, S, spec + (long)f * nlm, gx + f * ng, gy + f * ng, 1.0 / rsphere, ce,
                         ce2, four);
        free(S); free(ce); free(ce2); free(four);
    }
}

void orc_getvrtdivspec(const orc_sht *h, const double *u, const double *v, ocplx *vrt,
                       ocplx *div, int nfield, double rsphere) {
    long ng = (long)h->nlon * h->nlat;
    int nlm = h->nlm;
#pragma omp parallel
    {
        ocplx *ce = malloc(sizeof(ocplx) * h->nlme), *ce2 = malloc(sizeof(ocplx) * h->nlme);
        ocplx *four = malloc(sizeof(ocplx) * h->nlat * (h->T + 1));
#pragma omp for schedule(dynamic)
        for (int f = 0; f < nfield; f++) {
            grid_to_four(h, u + f * ng, four, 1);
            leg_anal(h, four, ce);
            grid_to_four(h, v + f * ng, four, 1);
            leg_anal(h, four, ce2);
            /* spat_to_sphtor then *lap/rsphere (shtns.f90:303,310) */
            uv_ext_to_vrtdiv(h, ce, ce2, vrt + (long)f * nlm, div + (long)f * nlm, rsphere);
        }
        free(ce); free(ce2); free(four);
    }
}
