// Column MLP (learned stand-in for the simple moist physics): a fully
// connected network nin -> hidden x nlayers (GELU) -> nout trained and run with
// hand-written bf16 MFMA GEMMs (FP32 accumulation), FP32 master weights and AdamW.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

namespace gfsdc {

struct MlpLayer {
    int K, N;       // real in features; out features padded to a multiple of 32
    int Nr;         // real out features
    int Kp;         // padded in features: K real, one bias column (= 1), zeros up to a multiple of 32
    float *w;       // FP32 master, [N][Kp] (bias in column K)
    float *m, *v;   // AdamW moments, [N][Kp]
    uint16_t *wt;   // bf16 copy [N][Kp] (forward operand)
    uint16_t *wk;   // bf16 copy [Kp][N] (input-gradient operand)
    float *gpart;   // split-batch partial gradients [S][N][Kp]
    // activations of the batch (input of this layer) and its gradient buffers
    uint16_t *a, *at;   // [Bp][Kp], [Kp][Bp] (bf16, transposed copy for the weight gradient)
    float *z;           // pre-activation of this layer's output [Bp][N] (FP32)
    uint16_t *dz, *dzt; // output gradient [Bp][N], [N][Bp] (bf16)
};

struct MlpNet {
    int nin = 0, nout = 0, hidden = 0, nlayers = 0;  // nlayers hidden layers -> nlayers + 1 GEMMs
    int maxbatch = 0, Bp = 0, S = 0;                 // Bp: batch rows padded to the tile; S: max batch splits
    int last_S = 1;                                  // splits used by the last backward
    std::vector<MlpLayer> L;
    int Np = 0;                 // padded outputs of the last layer
    float *y = nullptr;         // output [Bp][Np] FP32
    double *loss = nullptr;     // [1 + nblocks] loss accumulator
    int step = 0;               // AdamW step count (bias correction)
    uint64_t seed = 0;
};

int mlp_create(MlpNet **out, int nin, int nout, int hidden, int nlayers, int maxbatch, uint64_t seed,
               hipStream_t s);
void mlp_destroy(MlpNet *net);
// parameters in the interface layout: per layer W[K][N] (row-major, in x out), then b[N]
size_t mlp_nparams(const MlpNet *net);
int mlp_get_params(MlpNet *net, float *host, hipStream_t s);
int mlp_set_params(MlpNet *net, const float *host, hipStream_t s);
// forward of n rows of normalised input x (device FP32 [n][ldx]) into y (device FP32 [n][ldy])
int mlp_forward(MlpNet *net, const float *x, int ldx, int n, float *y, int ldy, hipStream_t s);
// forward + backward of one batch (n <= maxbatch) against normalised targets t
// (device FP32 [n][ldt]); MSE loss (mean over n * nout) to *loss_dev (device double)
int mlp_backward(MlpNet *net, const float *x, int ldx, const float *t, int ldt, int n, hipStream_t s);
// loss of the last mlp_backward (device double), gradients in the interface layout
const double *mlp_loss_dev(const MlpNet *net);
int mlp_get_grad(MlpNet *net, float *host, hipStream_t s);
// one AdamW update (beta 0.9 / 0.999, eps 1e-8, decoupled decay on weights, not biases)
// from the last mlp_backward, or from a host gradient in the interface layout
int mlp_adamw(MlpNet *net, double lr, double wd, const float *host_grad, hipStream_t s);

}  // namespace gfsdc
