/* FP64 restatement of the simple moist physics column (test infrastructure).
 *
 * Follows simple_physics_v5_1.f90:1-461 (Reed and Jablonowski 2012, version
 * 5.1): large-scale condensation (265-285), bulk surface fluxes of momentum,
 * sensible heat and moisture over a prescribed sea-surface temperature
 * (314-366) and implicit boundary-layer diffusion solved with a tridiagonal
 * sweep (375-447). Columns are ordered top -> bottom (index 1 = model top,
 * nl = lowest level), as in the reference routine. Pinned against the
 * reference routine itself, compiled from its source into oracle/_ref by
 * oracle/ref.mk, on random columns (tests/test_oracle_physics.py).
 *
 * sst_mode 0 / 1 are the reference's test = 0 (constant 302.15 K) and
 * test = 1 (moist baroclinic wave surface temperature, 231-239); sst_mode 2
 * (302.15 K - 30 K sin^2(lat)) has no counterpart in the reference.
 */
#include <math.h>
#include <stdlib.h>

#include "oracle.h"

#define SGL(x) ((double)(float)(x))

/* simple_physics_v5_1.f90:175-222 constants */
#define GRAVIT 9.80616
#define RAIR 287.0
#define CPAIR 1.0045e3
#define LATVAP 2.5e6
#define RH2O 461.5
#define C_SFC 0.0011
#define SST_TC 302.15
#define T0C 273.16
#define E0 610.78
#define RHOW 1000.0
#define CD0 0.0007
#define CD1 0.000065
#define CM 0.002
#define V20 20.0
#define P0 100000.0
#define PBLTOP 85000.0
#define PBLCONST 10000.0

double orc_simple_sst(int sst_mode, double lat) {
    if (sst_mode == 1) {
        /* 231-239: T00 = 288, u0 = 35, eta0 = 0.252, q0 = 0.021 (default-real literal) */
        const double pi = 4.0 * atan(1.0);
        const double a = 6371220.0, om = 7.29212e-5, zvir = (RH2O / RAIR) - 1.0;
        const double T00 = 288.0, u0 = 35.0, latw = 2.0 * pi / 9.0, eta0 = 0.252;
        const double etav = (1.0 - eta0) * 0.5 * pi, q0 = SGL(0.021);
        double s = sin(lat), c = cos(lat);
        double num = T00 + pi * u0 / RAIR * 1.5 * sin(etav) * pow(cos(etav), 0.5) *
                               ((-2.0 * pow(s, 6) * (c * c + 1.0 / 3.0) + 10.0 / 63.0) * u0 * pow(cos(etav), 1.5) +
                                (8.0 / 5.0 * c * c * c * (s * s + 2.0 / 3.0) - pi / 4.0) * a * om * 0.5);
        return num / (1.0 + zvir * q0 * exp(-pow(lat / latw, 4)));
    }
    if (sst_mode == 2) {
        double s = sin(lat);
        return SST_TC - 30.0 * s * s;
    }
    return SST_TC;
}

/* One column of simple_physics (pcols = 1). t, q, u, v updated in place;
 * *precl receives the large-scale precipitation rate (m/s). */
void orc_simple_column(int nl, double dtime, double lat, double *t, double *q, double *u, double *v,
                       const double *pmid, const double *pint, const double *pdel, const double *rpdel,
                       double ps, double *precl, int sst_mode, int turbflux) {
    const double epsilo = RAIR / RH2O, zvir = (RH2O / RAIR) - 1.0;
    /* 1-based views */
    double *T = t - 1, *Q = q - 1, *U = u - 1, *V = v - 1;
    const double *PM = pmid - 1, *PI = pint - 1, *PD = pdel - 1, *RPD = rpdel - 1;
    /* 225-229: height of the lowest full level, from the state before any update */
    double dlnpint = log(ps) - log(PI[nl]);
    double za = RAIR / GRAVIT * T[nl] * (1.0 + zvir * Q[nl]) * 0.5 * dlnpint;
    double tsurf = orc_simple_sst(sst_mode, lat);
    /* 265-292: large-scale condensation, then the state update */
    double pr = 0.0;
    for (int k = 1; k <= nl; k++) {
        double qsat = epsilo * E0 / PM[k] * exp(-LATVAP / RH2O * ((1.0 / T[k]) - 1.0 / T0C));
        if (Q[k] > qsat) {
            double tmp = 1.0 / dtime * (Q[k] - qsat) /
                         (1.0 + (LATVAP / CPAIR) * (epsilo * LATVAP * qsat / (RAIR * (T[k] * T[k]))));
            T[k] = T[k] + (LATVAP / CPAIR * tmp) * dtime;
            Q[k] = Q[k] + (-tmp) * dtime;
            pr = pr + tmp * PD[k] / (GRAVIT * RHOW);
        }
    }
    *precl = pr;
    if (!turbflux) return;
    /* 300-335: drag coefficients and eddy diffusivities */
    double wind = sqrt(U[nl] * U[nl] + V[nl] * V[nl]);
    double ke_s = C_SFC * wind * za, km_s, cd;
    if (wind < V20) {
        cd = CD0 + CD1 * wind;
        km_s = cd * wind * za;
    } else {
        cd = CM;
        km_s = CM * wind * za;
    }
    double *km = malloc(sizeof(double) * (nl + 2)), *ke = malloc(sizeof(double) * (nl + 2));
    for (int k = 1; k <= nl; k++) {
        if (PI[k] >= PBLTOP) {
            km[k] = km_s;
            ke[k] = ke_s;
        } else {
            double d = (PBLTOP - PI[k]) * (PBLTOP - PI[k]) / (PBLCONST * PBLCONST);
            km[k] = km_s * exp(-d);
            ke[k] = ke_s * exp(-d);
        }
    }
    km[nl + 1] = km_s;
    ke[nl + 1] = ke_s;
    /* 344-366: implicit surface fluxes on the lowest level */
    double qsats = epsilo * E0 / ps * exp(-LATVAP / RH2O * ((1.0 / tsurf) - 1.0 / T0C));
    double fm = 1.0 + cd * wind * dtime / za, fh = 1.0 + C_SFC * wind * dtime / za;
    U[nl] = U[nl] / fm;
    V[nl] = V[nl] / fm;
    T[nl] = (T[nl] + C_SFC * wind * tsurf * dtime / za) / fh;
    Q[nl] = (Q[nl] + C_SFC * wind * qsats * dtime / za) / fh;
    /* 375-447: tridiagonal boundary-layer diffusion (u, v on momentum
     * diffusivity; potential temperature and q on heat diffusivity) */
    int n2 = nl + 2;
    double *ca = calloc(n2, sizeof(double)), *cc = calloc(n2, sizeof(double));
    double *cam = calloc(n2, sizeof(double)), *ccm = calloc(n2, sizeof(double));
    double *ce = calloc(n2, sizeof(double)), *cem = calloc(n2, sizeof(double));
    double *cfu = calloc(n2, sizeof(double)), *cfv = calloc(n2, sizeof(double));
    double *cft = calloc(n2, sizeof(double)), *cfq = calloc(n2, sizeof(double));
    const double rk = RAIR / CPAIR;
    for (int k = 1; k <= nl - 1; k++) {
        double rho = (PI[k + 1] / (RAIR * (T[k + 1] + T[k]) / 2.0));
        double dpm = PM[k + 1] - PM[k];
        cam[k] = RPD[k] * dtime * GRAVIT * GRAVIT * km[k + 1] * rho * rho / dpm;
        ccm[k + 1] = RPD[k + 1] * dtime * GRAVIT * GRAVIT * km[k + 1] * rho * rho / dpm;
        ca[k] = RPD[k] * dtime * GRAVIT * GRAVIT * ke[k + 1] * rho * rho / dpm;
        cc[k + 1] = RPD[k + 1] * dtime * GRAVIT * GRAVIT * ke[k + 1] * rho * rho / dpm;
    }
    /* boundary values: ca(nl) = cc(1) = cam(nl) = ccm(1) = 0, ce(nl+1) = cf*(nl+1) = 0 (calloc) */
    for (int k = nl; k >= 1; k--) {
        double dh = 1.0 + ca[k] + cc[k] - ca[k] * ce[k + 1];
        double dm = 1.0 + cam[k] + ccm[k] - cam[k] * cem[k + 1];
        ce[k] = cc[k] / dh;
        cem[k] = ccm[k] / dm;
        cfu[k] = (U[k] + cam[k] * cfu[k + 1]) / dm;
        cfv[k] = (V[k] + cam[k] * cfv[k + 1]) / dm;
        cft[k] = (pow(P0 / PM[k], rk) * T[k] + ca[k] * cft[k + 1]) / dh;
        cfq[k] = (Q[k] + ca[k] * cfq[k + 1]) / dh;
    }
    U[1] = cfu[1];
    V[1] = cfv[1];
    T[1] = cft[1] * pow(PM[1] / P0, rk);
    Q[1] = cfq[1];
    for (int k = 2; k <= nl; k++) {
        U[k] = cem[k] * U[k - 1] + cfu[k];
        V[k] = cem[k] * V[k - 1] + cfv[k];
        T[k] = (ce[k] * T[k - 1] * pow(P0 / PM[k - 1], rk) + cft[k]) * pow(PM[k] / P0, rk);
        Q[k] = ce[k] * Q[k - 1] + cfq[k];
    }
    free(km); free(ke); free(ca); free(cc); free(cam); free(ccm); free(ce); free(cem);
    free(cfu); free(cfv); free(cft); free(cfq);
}
