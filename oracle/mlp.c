/* FP64 restatement of the column MLP of the learned physics (test
 * infrastructure): forward, mean-squared-error loss, backward and AdamW.
 * No counterpart in the reference; pinned by finite differences and the
 * AdamW closed form (tests/test_oracle_mlp.py).
 *
 * Parameters in one flat array per layer l = 1..L: W_l[K_l][N_l] (row-major,
 * input x output) then b_l[N_l]; K_1 = nin, N_L = nout, hidden widths between;
 * GELU (erf form) after every layer but the last. */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static void dims(int l, int nin, int nout, int hidden, int nlayers, int *K, int *N) {
    *K = (l == 0) ? nin : hidden;
    *N = (l == nlayers) ? nout : hidden;
}

long orc_mlp_nparams(int nin, int nout, int hidden, int nlayers) {
    long n = 0;
    for (int l = 0; l <= nlayers; l++) {
        int K, N;
        dims(l, nin, nout, hidden, nlayers, &K, &N);
        n += (long)K * N + N;
    }
    return n;
}

static double gelu(double x) { return 0.5 * x * (1.0 + erf(x / sqrt(2.0))); }
static double dgelu(double x) { return 0.5 * (1.0 + erf(x / sqrt(2.0))) + x * exp(-0.5 * x * x) / sqrt(2.0 * M_PI); }

/* value rounded to bfloat16 (through float, each step round-to-nearest-even) */
static double bf(double v) {
    float f = (float)v;
    uint32_t u;
    memcpy(&u, &f, 4);
    u = (u + 0x7FFFu + ((u >> 16) & 1u)) & 0xFFFF0000u;
    memcpy(&f, &u, 4);
    return (double)f;
}

/* loss = mean over n*nout of (y - t)^2; grad (may be NULL) = dloss/dparams;
 * y (may be NULL) receives the network output [n][nout]. With rb != 0 the
 * network is evaluated with the storage rounding of a bf16 matrix-core
 * network: inputs, weights and biases, each hidden activation (after GELU),
 * the output gradient and each back-propagated pre-activation gradient are
 * rounded to bfloat16; sums stay FP64 and the output layer is not rounded. */
static double mlp_run(int nin, int nout, int hidden, int nlayers, const double *params, int n,
                      const double *x, const double *t, double *grad, double *y, int rb) {
    const int nL = nlayers + 1;
    double **act = malloc(sizeof(double *) * (nL + 1)); /* act[l]: input of layer l, [n][K_l] */
    double **pre = malloc(sizeof(double *) * nL);       /* pre-activation of layer l output */
    const double **W = malloc(sizeof(double *) * nL);
    long *off = malloc(sizeof(long) * nL);
    long o = 0;
    for (int l = 0; l < nL; l++) {
        int K, N;
        dims(l, nin, nout, hidden, nlayers, &K, &N);
        off[l] = o;
        W[l] = params + o;
        o += (long)K * N + N;
    }
    double *wr = NULL; /* rounded copy of the parameters */
    if (rb) {
        wr = malloc(sizeof(double) * o);
        for (long i = 0; i < o; i++) wr[i] = bf(params[i]);
        for (int l = 0; l < nL; l++) W[l] = wr + off[l];
    }
    act[0] = malloc(sizeof(double) * n * nin);
    for (long i = 0; i < (long)n * nin; i++) act[0][i] = rb ? bf(x[i]) : x[i];
    for (int l = 0; l < nL; l++) {
        int K, N;
        dims(l, nin, nout, hidden, nlayers, &K, &N);
        const double *w = W[l], *b = W[l] + (long)K * N;
        pre[l] = malloc(sizeof(double) * n * N);
        act[l + 1] = malloc(sizeof(double) * n * N);
        for (int r = 0; r < n; r++)
            for (int j = 0; j < N; j++) {
                double s = b[j];
                for (int k = 0; k < K; k++) s += act[l][(long)r * K + k] * w[(long)k * N + j];
                pre[l][(long)r * N + j] = s;
                act[l + 1][(long)r * N + j] = (l < nL - 1) ? (rb ? bf(gelu(s)) : gelu(s)) : s;
            }
    }
    double loss = 0.0;
    const double *out = act[nL];
    for (long i = 0; i < (long)n * nout; i++) loss += (out[i] - t[i]) * (out[i] - t[i]);
    loss /= (double)n * nout;
    if (y) memcpy(y, out, sizeof(double) * n * nout);
    if (grad) {
        memset(grad, 0, sizeof(double) * o);
        double *d = malloc(sizeof(double) * n * nout); /* d loss / d pre of the current layer */
        for (long i = 0; i < (long)n * nout; i++) {
            d[i] = 2.0 * (out[i] - t[i]) / ((double)n * nout);
            if (rb) d[i] = bf(d[i]);
        }
        for (int l = nL - 1; l >= 0; l--) {
            int K, N;
            dims(l, nin, nout, hidden, nlayers, &K, &N);
            double *gw = grad + off[l], *gb = gw + (long)K * N;
            for (int r = 0; r < n; r++)
                for (int j = 0; j < N; j++) {
                    double dj = d[(long)r * N + j];
                    gb[j] += dj;
                    for (int k = 0; k < K; k++) gw[(long)k * N + j] += act[l][(long)r * K + k] * dj;
                }
            if (l > 0) {
                double *dp = malloc(sizeof(double) * n * K);
                const double *w = W[l];
                for (int r = 0; r < n; r++)
                    for (int k = 0; k < K; k++) {
                        double s = 0.0;
                        for (int j = 0; j < N; j++) s += d[(long)r * N + j] * w[(long)k * N + j];
                        dp[(long)r * K + k] = s * dgelu(pre[l - 1][(long)r * K + k]);
                        if (rb) dp[(long)r * K + k] = bf(dp[(long)r * K + k]);
                    }
                free(d);
                d = dp;
            }
        }
        free(d);
    }
    for (int l = 0; l <= nL; l++) free(act[l]);
    for (int l = 0; l < nL; l++) free(pre[l]);
    free(act); free(pre); free(W); free(off); free(wr);
    return loss;
}

double orc_mlp_loss_grad(int nin, int nout, int hidden, int nlayers, const double *params, int n,
                         const double *x, const double *t, double *grad, double *y) {
    return mlp_run(nin, nout, hidden, nlayers, params, n, x, t, grad, y, 0);
}

double orc_mlp_loss_grad_bf16(int nin, int nout, int hidden, int nlayers, const double *params, int n,
                              const double *x, const double *t, double *grad, double *y) {
    return mlp_run(nin, nout, hidden, nlayers, params, n, x, t, grad, y, 1);
}

/* one AdamW step (beta1 0.9, beta2 0.999, eps 1e-8; decoupled decay on the
 * weights, not the biases) at step count `step` (1-based) */
void orc_mlp_adamw(int nin, int nout, int hidden, int nlayers, double *params, double *m, double *v,
                   const double *grad, int step, double lr, double wd) {
    const double b1 = 0.9, b2 = 0.999, eps = 1e-8;
    const double c1 = 1.0 - pow(b1, step), c2 = 1.0 - pow(b2, step);
    long o = 0;
    for (int l = 0; l <= nlayers; l++) {
        int K, N;
        dims(l, nin, nout, hidden, nlayers, &K, &N);
        long nw = (long)K * N, np = nw + N;
        for (long i = 0; i < np; i++) {
            long q = o + i;
            m[q] = b1 * m[q] + (1.0 - b1) * grad[q];
            v[q] = b2 * v[q] + (1.0 - b2) * grad[q] * grad[q];
            double p = params[q];
            if (i < nw) p *= 1.0 - lr * wd;
            params[q] = p - lr * (m[q] / c1) / (sqrt(v[q] / c2) + eps);
        }
        o += np;
    }
}
