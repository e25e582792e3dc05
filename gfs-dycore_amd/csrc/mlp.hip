// Column MLP on CDNA4: bf16 MFMA (v_mfma_f32_16x16x32_bf16) GEMMs with FP32
// accumulation and fused epilogues, FP32 master weights, AdamW.
//
// Every GEMM is C[m][n] = sum_k A[m][k] B[n][k] with both operands stored
// k-contiguous (bf16), so each lane feeds the matrix core 8 consecutive k
// values of one row of A and one row of B with a single 16-byte LDS read.
//   forward   z_l   = a_l  . wt_l^T        a_l [B][Kp], wt_l [N][Kp]
//   input grad da_l = dz_l . wk_l^T        dz_l [B][N], wk_l [Kp][N]  (rows k < K)
//   weight grad dW_l^T = dzt_l . at_l^T    dzt_l [N][B], at_l [Kp][B] (split over B)
// The bias is the weight of a constant-one input column (index K of Kp), so
// bias gradients come out of the weight-gradient GEMM. Activation epilogues
// write both the row-major and the transposed bf16 copy the next GEMMs read.
#include "mlp.h"

#include <algorithm>
#include <cmath>
#include <cstring>

#include "engine.h"

namespace gfsdc {

typedef unsigned short us8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int TM = 128, TN = 128, TK = 32, LPAD = 8;

__host__ __device__ inline uint16_t f2bf(float f) {
    // round to nearest even (finite inputs)
    uint32_t u = __builtin_bit_cast(uint32_t, f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__device__ inline float gelu(float x) { return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f)); }
// forward GELU with erf from Abramowitz & Stegun (1964) 7.1.25 (|error| <= 2.5e-5,
// below the bf16 rounding of the activations, 2^-9) on the hardware exp2 / rcp
__device__ inline float gelu_fast(float x) {
    const float z = x * 0.70710678118654752f, a = fabsf(z);
    const float t = __builtin_amdgcn_rcpf(1.0f + 0.47047f * a);
    const float p = t * (0.3480242f + t * (-0.0958798f + t * 0.7478556f));
    const float y = 1.0f - p * __builtin_amdgcn_exp2f(-a * a * 1.4426950408889634f);
    return 0.5f * x * (1.0f + copysignf(y, z));
}
// two floats to packed bf16 (round to nearest even, the hardware conversion)
__device__ inline uint32_t pack_bf16(float a, float b) {
    return (uint32_t)__builtin_bit_cast(uint16_t, (__bf16)a) | ((uint32_t)__builtin_bit_cast(uint16_t, (__bf16)b) << 16);
}
__device__ inline float dgelu(float x) {
    return 0.5f * (1.0f + erff(x * 0.70710678118654752f)) + x * 0.39894228040143268f * expf(-0.5f * x * x);
}

enum { EPI_F32 = 0, EPI_GELU = 1, EPI_DGELU = 2 };

struct Epi {
    float *c;       // EPI_F32 output, c[z * cslice + m * ldc + n]
    long ldc, cslice;
    float *z;       // EPI_GELU: pre-activation written; EPI_DGELU: read
    long ldz;
    uint16_t *o;    // bf16 [m][n]
    long ldo;
    uint16_t *ot;   // bf16 [n][m]
    long ldot;
};

// 128 x 128 output tile per workgroup (4 waves of 64 x 64 = 4 x 4 MFMA blocks),
// k in chunks of 32 staged through LDS; blockIdx.z selects a k slice of length Kc.
// SWAP multiplies B A^T instead of A B^T, so a lane's four accumulators are four
// consecutive columns of one output row: vector stores of the forward outputs
// (the training forward, which also writes the transposed activations, keeps
// four consecutive rows per lane instead).
template <int EPI, bool SWAP = false>
__global__ __launch_bounds__(256) void k_gemm_nt(const uint16_t *__restrict__ A, long lda,
                                                 const uint16_t *__restrict__ B, long ldb, int M, int N, int Kc,
                                                 Epi e) {
    // two LDS stages: chunk k+1 is stored into the other stage while chunk k is
    // multiplied, so one barrier per chunk suffices
    __shared__ __attribute__((aligned(16))) uint16_t As[2][TM][TK + LPAD];
    __shared__ __attribute__((aligned(16))) uint16_t Bs[2][TN][TK + LPAD];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int m0 = blockIdx.y * TM, n0 = blockIdx.x * TN;
    const long k0 = (long)blockIdx.z * Kc;
    const int wm = (wave >> 1) * 64, wn = (wave & 1) * 64;
    f4 acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};
    const us8 zero = {0, 0, 0, 0, 0, 0, 0, 0};
    us8 ra[2], rb[2];
    auto load = [&](long kk) {
#pragma unroll
        for (int u = 0; u < 2; u++) {
            const int c = tid + 256 * u, r = c >> 2, q = (c & 3) * 8;
            ra[u] = (m0 + r < M) ? *(const us8 *)(A + (long)(m0 + r) * lda + kk + q) : zero;
            rb[u] = (n0 + r < N) ? *(const us8 *)(B + (long)(n0 + r) * ldb + kk + q) : zero;
        }
    };
    auto stage = [&](int b) {
#pragma unroll
        for (int u = 0; u < 2; u++) {
            const int c = tid + 256 * u, r = c >> 2, q = (c & 3) * 8;
            *(us8 *)&As[b][r][q] = ra[u];
            *(us8 *)&Bs[b][r][q] = rb[u];
        }
    };
    load(k0);
    stage(0);
    __syncthreads();
    int b = 0;
    for (long kk = k0; kk < k0 + Kc; kk += TK) {
        const bool next = kk + TK < k0 + Kc;
        if (next) load(kk + TK);
        bf16x8 fa[4], fb[4];
#pragma unroll
        for (int i = 0; i < 4; i++) {
            fa[i] = __builtin_bit_cast(bf16x8, *(const us8 *)&As[b][wm + i * 16 + (lane & 15)][8 * (lane >> 4)]);
            fb[i] = __builtin_bit_cast(bf16x8, *(const us8 *)&Bs[b][wn + i * 16 + (lane & 15)][8 * (lane >> 4)]);
        }
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int j = 0; j < 4; j++) {
                if constexpr (SWAP)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fb[j], fa[i], acc[i][j], 0, 0, 0);
                else
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[i], fb[j], acc[i][j], 0, 0, 0);
            }
        // the other stage was last read before the previous barrier: free to fill
        if (next) stage(b ^ 1);
        __syncthreads();
        b ^= 1;
    }
    if constexpr (SWAP) {
        // lane holds row lane % 16, columns 4 (lane / 16) + r of each 16 x 16 block
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int m = m0 + wm + i * 16 + (lane & 15);
            if (m >= M) continue;
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int n = n0 + wn + j * 16 + 4 * (lane >> 4);  // N is a multiple of 32: n + 3 < N
                if (n >= N) continue;
                const f4 v = acc[i][j];
                if constexpr (EPI == EPI_F32) {
                    *(float4 *)(e.c + (long)m * e.ldc + n) = float4{v[0], v[1], v[2], v[3]};
                } else {
                    if (e.z) *(float4 *)(e.z + (long)m * e.ldz + n) = float4{v[0], v[1], v[2], v[3]};
                    uint2 pk;
                    pk.x = pack_bf16(gelu_fast(v[0]), gelu_fast(v[1]));
                    pk.y = pack_bf16(gelu_fast(v[2]), gelu_fast(v[3]));
                    *(uint2 *)(e.o + (long)m * e.ldo + n) = pk;
                }
            }
        }
        return;
    }
    // lane holds rows 4 (lane / 16) + r, column lane % 16 of each 16 x 16 block
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int mb = m0 + wm + i * 16 + 4 * (lane >> 4), n = n0 + wn + j * 16 + (lane & 15);
            if (n >= N) continue;
            if (EPI == EPI_F32) {
                float *c = e.c + (long)blockIdx.z * e.cslice;
#pragma unroll
                for (int r = 0; r < 4; r++)
                    if (mb + r < M) c[(long)(mb + r) * e.ldc + n] = acc[i][j][r];
            } else {
                uint16_t h[4];
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int m = mb + r;
                    float v = acc[i][j][r];
                    if (m < M) {
                        if (EPI == EPI_GELU) {
                            if (e.z) e.z[(long)m * e.ldz + n] = v;  // kept for the backward pass only
                            v = gelu_fast(v);
                        } else {
                            v *= dgelu(e.z[(long)m * e.ldz + n]);
                        }
                        h[r] = f2bf(v);
                        e.o[(long)m * e.ldo + n] = h[r];
                    } else {
                        h[r] = 0;
                    }
                }
                // four consecutive rows of one column: one 8-byte store in the transposed copy
                // (read by the weight-gradient GEMM; inference passes none)
                if (!e.ot) {
                } else if (mb + 3 < M) {
                    uint2 pk;
                    pk.x = (uint32_t)h[0] | ((uint32_t)h[1] << 16);
                    pk.y = (uint32_t)h[2] | ((uint32_t)h[3] << 16);
                    *(uint2 *)(e.ot + (long)n * e.ldot + mb) = pk;
                } else {
                    for (int r = 0; r < 4; r++)
                        if (mb + r < M) e.ot[(long)n * e.ldot + mb + r] = h[r];
                }
            }
        }
}

// network input: x (FP32 [n][ldx], nin features) -> a0 [Bp][Kp], at0 [Kp][Bp] with
// the constant-one bias column at nin and zero padding
__global__ void k_act_in(const float *__restrict__ x, int ldx, int n, int nin, int Kp, int Mb, int Bp,
                         uint16_t *__restrict__ a, uint16_t *__restrict__ at) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)Mb * Kp) return;
    const int b = (int)(i / Kp), k = (int)(i % Kp);
    float v = 0.f;
    if (k == nin) v = 1.f;
    else if (k < nin && b < n) v = x[(long)b * ldx + k];
    const uint16_t h = f2bf(v);
    a[i] = h;
    if (at) at[(long)k * Bp + b] = h;  // transposed copy for the weight gradient (training only)
}

// constant parts of a hidden activation buffer: bias column = 1, padding = 0
__global__ void k_act_init(uint16_t *__restrict__ a, uint16_t *__restrict__ at, int K, int Kp, int Bp) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)Bp * Kp) return;
    const int b = (int)(i / Kp), k = (int)(i % Kp);
    const uint16_t h = (k == K) ? f2bf(1.f) : (uint16_t)0;
    a[i] = h;
    at[(long)k * Bp + b] = h;
}

// MSE loss over n rows x nout outputs; gradient 2 (y - t) / (n nout) into dz/dzt
// (zero on padded rows); per-block loss partial sums
__global__ __launch_bounds__(256) void k_loss(const float *__restrict__ y, int Np, const float *__restrict__ t,
                                              int ldt, int n, int nout, int Mb, int Bp, uint16_t *__restrict__ dz,
                                              uint16_t *__restrict__ dzt, double *__restrict__ part) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    double l = 0.0;
    if (i < (long)Mb * Np) {
        const int b = (int)(i / Np), j = (int)(i % Np);
        float g = 0.f;
        if (b < n && j < nout) {
            const float d = y[(long)b * Np + j] - t[(long)b * ldt + j];
            l = (double)d * d;
            g = 2.f * d / ((float)n * (float)nout);
        }
        const uint16_t h = f2bf(g);
        dz[i] = h;
        dzt[(long)j * Bp + b] = h;
    }
    __shared__ double red[256];
    red[threadIdx.x] = l;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[1 + blockIdx.x] = red[0];
}

__global__ __launch_bounds__(256) void k_loss_final(double *__restrict__ part, int nblk, double scale) {
    __shared__ double red[256];
    double s = 0.0;
    for (int i = threadIdx.x; i < nblk; i += 256) s += part[1 + i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[0] = red[0] * scale;
}

// AdamW on one layer's [N][Kp] parameters (column K = bias: no decay; columns
// beyond K stay 0); gradient = sum of S split partials or a given array.
// Refreshes both bf16 copies.
__global__ void k_adamw(float *__restrict__ w, float *__restrict__ m, float *__restrict__ v,
                        const float *__restrict__ gpart, int S, const float *__restrict__ gin, int N, int K,
                        int Kp, float lr, float wd, float b1c, float b2c, uint16_t *__restrict__ wt,
                        uint16_t *__restrict__ wk) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long np = (long)N * Kp;
    if (i >= np) return;
    const int n = (int)(i / Kp), k = (int)(i % Kp);
    float p = w[i];
    if (k <= K) {
        float g;
        if (gin) {
            g = gin[i];
        } else {
            g = 0.f;
            for (int s = 0; s < S; s++) g += gpart[s * np + i];
        }
        const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
        const float mm = b1 * m[i] + (1.f - b1) * g;
        const float vv = b2 * v[i] + (1.f - b2) * g * g;
        m[i] = mm;
        v[i] = vv;
        if (k < K) p = p * (1.f - lr * wd);
        p = p - lr * (mm / b1c) / (sqrtf(vv / b2c) + eps);
        w[i] = p;
    }
    const uint16_t h = f2bf(p);
    wt[i] = h;
    wk[(long)k * N + n] = h;
}

// refresh the bf16 copies from the master
__global__ void k_wcopy(const float *__restrict__ w, int N, int Kp, uint16_t *__restrict__ wt,
                        uint16_t *__restrict__ wk) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)N * Kp) return;
    const int n = (int)(i / Kp), k = (int)(i % Kp);
    const uint16_t h = f2bf(w[i]);
    wt[i] = h;
    wk[(long)k * N + n] = h;
}

// ---------------------------------------------------------------------------
// host side


template <typename X> static X *dal(size_t n) {
    X *p = nullptr;
    if (hipMalloc((void **)&p, sizeof(X) * (n ? n : 1)) != hipSuccess) return nullptr;
    return p;
}

static uint64_t mix64(uint64_t seed, uint64_t stream, uint64_t index) {
    uint64_t z = seed * 0x9E3779B97F4A7C15ULL ^ (stream + 0x632BE59BD9B4E019ULL) * 0xBF58476D1CE4E5B9ULL;
    z += index * 0x94D049BB133111EBULL + 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

int mlp_create(MlpNet **out, int nin, int nout, int hidden, int nlayers, int maxbatch, uint64_t seed,
               hipStream_t s) {
    *out = nullptr;
    if (nin < 1 || nout < 1 || hidden < 32 || hidden % 32 || nlayers < 1 || maxbatch < 1) {
        set_error("mlp: need nin, nout >= 1, hidden a multiple of 32, nlayers >= 1, maxbatch >= 1");
        return GFSDC_EINVAL;
    }
    MlpNet *net = new MlpNet();
    net->nin = nin;
    net->nout = nout;
    net->hidden = hidden;
    net->nlayers = nlayers;
    net->seed = seed;
    net->Bp = round_up(maxbatch, TM);
    net->maxbatch = net->Bp;
    net->S = 32;  // capacity of the split-batch gradient partials
    const int Bp = net->Bp, nL = nlayers + 1;
    net->L.resize(nL);
    bool ok = true;
    for (int l = 0; l < nL; l++) {
        MlpLayer &Ly = net->L[l];
        Ly.K = (l == 0) ? nin : hidden;
        Ly.Nr = (l == nL - 1) ? nout : hidden;
        Ly.N = round_up(Ly.Nr, TK);  // padded outputs carry zero weights and no loss
        Ly.Kp = round_up(Ly.K + 1, TK);
        const size_t np = (size_t)Ly.N * Ly.Kp;
        Ly.w = dal<float>(np);
        Ly.m = dal<float>(np);
        Ly.v = dal<float>(np);
        Ly.wt = dal<uint16_t>(np);
        Ly.wk = dal<uint16_t>(np);
        Ly.gpart = dal<float>(np * net->S);
        Ly.a = dal<uint16_t>((size_t)Bp * Ly.Kp);
        Ly.at = dal<uint16_t>((size_t)Bp * Ly.Kp);
        Ly.z = dal<float>((size_t)Bp * Ly.N);
        Ly.dz = dal<uint16_t>((size_t)Bp * Ly.N);
        Ly.dzt = dal<uint16_t>((size_t)Bp * Ly.N);
        ok = ok && Ly.w && Ly.m && Ly.v && Ly.wt && Ly.wk && Ly.gpart && Ly.a && Ly.at && Ly.z && Ly.dz && Ly.dzt;
    }
    net->Np = net->L[nL - 1].N;
    net->y = dal<float>((size_t)Bp * net->Np);
    const int nblk = (int)(((long)Bp * net->Np + 255) / 256);
    net->loss = dal<double>(1 + nblk);
    if (!ok || !net->y || !net->loss) {
        mlp_destroy(net);
        set_error("mlp: device allocation failed");
        return GFSDC_ENOMEM;
    }
    // Xavier-uniform weights from the counter RNG (seed, stream = layer, index = n*K+k), zero biases
    for (int l = 0; l < nL; l++) {
        MlpLayer &Ly = net->L[l];
        std::vector<float> w((size_t)Ly.N * Ly.Kp, 0.f);
        const double a = std::sqrt(6.0 / (Ly.K + Ly.Nr));
        for (int n = 0; n < Ly.Nr; n++)
            for (int k = 0; k < Ly.K; k++) {
                double u = (double)(mix64(seed, 0x4D4C50ULL + l, (uint64_t)n * Ly.K + k) >> 11) * 0x1p-53;
                w[(size_t)n * Ly.Kp + k] = (float)((2.0 * u - 1.0) * a);
            }
        const size_t np = w.size();
        GFSDC_CHECK(hipMemcpyAsync(Ly.w, w.data(), sizeof(float) * np, hipMemcpyHostToDevice, s));
        GFSDC_CHECK(hipMemsetAsync(Ly.m, 0, sizeof(float) * np, s));
        GFSDC_CHECK(hipMemsetAsync(Ly.v, 0, sizeof(float) * np, s));
        hipLaunchKernelGGL(k_wcopy, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, s, Ly.w, Ly.N, Ly.Kp, Ly.wt,
                           Ly.wk);
        if (l > 0) {
            const long na = (long)Bp * Ly.Kp;
            hipLaunchKernelGGL(k_act_init, dim3((unsigned)((na + 255) / 256)), dim3(256), 0, s, Ly.a, Ly.at, Ly.K,
                               Ly.Kp, Bp);
        }
        // the host vector must outlive the copy
        GFSDC_CHECK(hipStreamSynchronize(s));
    }
    GFSDC_CHECK(hipGetLastError());
    *out = net;
    return 0;
}

void mlp_destroy(MlpNet *net) {
    if (!net) return;
    for (auto &Ly : net->L) {
        void *p[] = {Ly.w, Ly.m, Ly.v, Ly.wt, Ly.wk, Ly.gpart, Ly.a, Ly.at, Ly.z, Ly.dz, Ly.dzt};
        for (void *q : p)
            if (q) (void)hipFree(q);
    }
    if (net->y) (void)hipFree(net->y);
    if (net->loss) (void)hipFree(net->loss);
    delete net;
}

size_t mlp_nparams(const MlpNet *net) {
    size_t n = 0;
    for (auto &Ly : net->L) n += (size_t)Ly.K * Ly.Nr + Ly.Nr;
    return n;
}

// interface layout (W[K][N], b[N]) <-> device layout [N][Kp] with the bias in column K
static void to_device_layout(const MlpLayer &Ly, const float *src, std::vector<float> &dst) {
    dst.assign((size_t)Ly.N * Ly.Kp, 0.f);
    for (int k = 0; k < Ly.K; k++)
        for (int n = 0; n < Ly.Nr; n++) dst[(size_t)n * Ly.Kp + k] = src[(size_t)k * Ly.Nr + n];
    for (int n = 0; n < Ly.Nr; n++) dst[(size_t)n * Ly.Kp + Ly.K] = src[(size_t)Ly.K * Ly.Nr + n];
}
static void from_device_layout(const MlpLayer &Ly, const std::vector<float> &src, float *dst) {
    for (int k = 0; k < Ly.K; k++)
        for (int n = 0; n < Ly.Nr; n++) dst[(size_t)k * Ly.Nr + n] = src[(size_t)n * Ly.Kp + k];
    for (int n = 0; n < Ly.Nr; n++) dst[(size_t)Ly.K * Ly.Nr + n] = src[(size_t)n * Ly.Kp + Ly.K];
}

int mlp_get_params(MlpNet *net, float *host, hipStream_t s) {
    size_t off = 0;
    for (auto &Ly : net->L) {
        std::vector<float> d((size_t)Ly.N * Ly.Kp);
        GFSDC_CHECK(hipMemcpyAsync(d.data(), Ly.w, sizeof(float) * d.size(), hipMemcpyDeviceToHost, s));
        GFSDC_CHECK(hipStreamSynchronize(s));
        from_device_layout(Ly, d, host + off);
        off += (size_t)Ly.K * Ly.Nr + Ly.Nr;
    }
    return 0;
}

int mlp_set_params(MlpNet *net, const float *host, hipStream_t s) {
    size_t off = 0;
    for (auto &Ly : net->L) {
        std::vector<float> d;
        to_device_layout(Ly, host + off, d);
        GFSDC_CHECK(hipMemcpyAsync(Ly.w, d.data(), sizeof(float) * d.size(), hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(k_wcopy, dim3((unsigned)((d.size() + 255) / 256)), dim3(256), 0, s, Ly.w, Ly.N, Ly.Kp,
                           Ly.wt, Ly.wk);
        GFSDC_CHECK(hipStreamSynchronize(s));
        off += (size_t)Ly.K * Ly.Nr + Ly.Nr;
    }
    net->step = 0;
    for (auto &Ly : net->L) {
        GFSDC_CHECK(hipMemsetAsync(Ly.m, 0, sizeof(float) * Ly.N * Ly.Kp, s));
        GFSDC_CHECK(hipMemsetAsync(Ly.v, 0, sizeof(float) * Ly.N * Ly.Kp, s));
    }
    return 0;
}

// batch splits of the weight-gradient GEMM for Mb rows: slices of >= 256 rows,
// a multiple of the k chunk
static int splits(const MlpNet *net, int Mb) {
    int S = std::max(1, std::min(net->S, Mb / 256));
    while (Mb % (S * TK)) S--;
    return S;
}

// forward of Mb (multiple of TM) rows already in L[0].a / at; output in net->y.
// Training keeps the pre-activations and the transposed activations for the
// backward pass; inference writes only what the next layer reads.
// out / ldout / nrow: where the last layer writes (default net->y, all Mb rows);
// inference can hand the caller's rows straight to the output GEMM.
static void forward_batch(MlpNet *net, int Mb, bool train, hipStream_t s, float *out = nullptr, long ldout = 0,
                          int nrow = 0) {
    const int nL = (int)net->L.size(), Bp = net->Bp;
    for (int l = 0; l < nL; l++) {
        MlpLayer &Ly = net->L[l];
        dim3 g((unsigned)((Ly.N + TN - 1) / TN), (unsigned)(Mb / TM), 1);
        Epi e{};
        if (l == nL - 1) {
            e.c = out ? out : net->y;
            e.ldc = out ? ldout : Ly.N;  // = Np
            hipLaunchKernelGGL((k_gemm_nt<EPI_F32, true>), g, dim3(256), 0, s, Ly.a, (long)Ly.Kp, Ly.wt, (long)Ly.Kp,
                               out ? nrow : Mb, Ly.N, Ly.Kp, e);
        } else {
            MlpLayer &Nx = net->L[l + 1];
            e.z = train ? Ly.z : nullptr;
            e.ldz = Ly.N;
            e.o = Nx.a;
            e.ldo = Nx.Kp;
            e.ot = train ? Nx.at : nullptr;
            e.ldot = Bp;
            if (train)
                hipLaunchKernelGGL(k_gemm_nt<EPI_GELU>, g, dim3(256), 0, s, Ly.a, (long)Ly.Kp, Ly.wt, (long)Ly.Kp,
                                   Mb, Ly.N, Ly.Kp, e);
            else
                hipLaunchKernelGGL((k_gemm_nt<EPI_GELU, true>), g, dim3(256), 0, s, Ly.a, (long)Ly.Kp, Ly.wt,
                                   (long)Ly.Kp, Mb, Ly.N, Ly.Kp, e);
        }
    }
}

static int input_batch(MlpNet *net, const float *x, int ldx, int n, bool train, hipStream_t s) {
    MlpLayer &L0 = net->L[0];
    const int Mb = round_up(n, TM);
    const long na = (long)Mb * L0.Kp;
    hipLaunchKernelGGL(k_act_in, dim3((unsigned)((na + 255) / 256)), dim3(256), 0, s, x, ldx, n, net->nin, L0.Kp,
                       Mb, net->Bp, L0.a, train ? L0.at : (uint16_t *)nullptr);
    return Mb;
}

int mlp_forward(MlpNet *net, const float *x, int ldx, int n, float *y, int ldy, hipStream_t s) {
    for (int r0 = 0; r0 < n; r0 += net->Bp) {
        const int nb = std::min(net->Bp, n - r0);
        const int Mb = input_batch(net, x + (long)r0 * ldx, ldx, nb, false, s);
        if (net->Np == net->nout && ldy >= net->Np && ldy % 4 == 0 && ((uintptr_t)y & 15) == 0) {
            // no padded outputs: the last GEMM writes the nb rows of y itself
            forward_batch(net, Mb, false, s, y + (long)r0 * ldy, ldy, nb);
        } else {
            forward_batch(net, Mb, false, s);
            GFSDC_CHECK(hipMemcpy2DAsync(y + (long)r0 * ldy, sizeof(float) * ldy, net->y, sizeof(float) * net->Np,
                                         sizeof(float) * net->nout, nb, hipMemcpyDeviceToDevice, s));
        }
    }
    GFSDC_CHECK(hipGetLastError());
    return 0;
}

int mlp_backward(MlpNet *net, const float *x, int ldx, const float *t, int ldt, int n, hipStream_t s) {
    if (n > net->Bp) {
        set_error("mlp: batch larger than maxbatch");
        return GFSDC_EINVAL;
    }
    const int nL = (int)net->L.size(), Bp = net->Bp;
    const int Mb = input_batch(net, x, ldx, n, true, s);
    const int S = splits(net, Mb);
    net->last_S = S;
    forward_batch(net, Mb, true, s);
    MlpLayer &Lo = net->L[nL - 1];
    const int nblk = (int)(((long)Mb * net->Np + 255) / 256);
    hipLaunchKernelGGL(k_loss, dim3((unsigned)nblk), dim3(256), 0, s, net->y, net->Np, t, ldt, n, net->nout, Mb,
                       Bp, Lo.dz, Lo.dzt, net->loss);
    hipLaunchKernelGGL(k_loss_final, dim3(1), dim3(256), 0, s, net->loss, nblk,
                       1.0 / ((double)n * net->nout));
    for (int l = nL - 1; l >= 0; l--) {
        MlpLayer &Ly = net->L[l];
        // weight gradient (transposed, bias column included), split over the batch
        {
            Epi e{};
            e.c = Ly.gpart;
            e.ldc = Ly.Kp;
            e.cslice = (long)Ly.N * Ly.Kp;
            dim3 g((unsigned)((Ly.Kp + TN - 1) / TN), (unsigned)((Ly.N + TM - 1) / TM), (unsigned)S);
            hipLaunchKernelGGL(k_gemm_nt<EPI_F32>, g, dim3(256), 0, s, Ly.dzt, (long)Bp, Ly.at, (long)Bp, Ly.N,
                               Ly.Kp, Mb / S, e);
        }
        if (l > 0) {
            // input gradient through the GELU of the layer below
            MlpLayer &Pv = net->L[l - 1];
            Epi e{};
            e.z = Pv.z;
            e.ldz = Pv.N;
            e.o = Pv.dz;
            e.ldo = Pv.N;
            e.ot = Pv.dzt;
            e.ldot = Bp;
            dim3 g((unsigned)((Ly.K + TN - 1) / TN), (unsigned)(Mb / TM), 1);
            hipLaunchKernelGGL(k_gemm_nt<EPI_DGELU>, g, dim3(256), 0, s, Ly.dz, (long)Ly.N, Ly.wk, (long)Ly.N, Mb,
                               Ly.K, Ly.N, e);
        }
    }
    GFSDC_CHECK(hipGetLastError());
    return 0;
}

const double *mlp_loss_dev(const MlpNet *net) { return net->loss; }

int mlp_get_grad(MlpNet *net, float *host, hipStream_t s) {
    size_t off = 0;
    for (auto &Ly : net->L) {
        const size_t np = (size_t)Ly.N * Ly.Kp;
        std::vector<float> p(np * net->last_S), d(np, 0.f);
        GFSDC_CHECK(hipMemcpyAsync(p.data(), Ly.gpart, sizeof(float) * p.size(), hipMemcpyDeviceToHost, s));
        GFSDC_CHECK(hipStreamSynchronize(s));
        for (int q = 0; q < net->last_S; q++)
            for (size_t i = 0; i < np; i++) d[i] += p[q * np + i];
        from_device_layout(Ly, d, host + off);
        off += (size_t)Ly.K * Ly.Nr + Ly.Nr;
    }
    return 0;
}

int mlp_adamw(MlpNet *net, double lr, double wd, const float *host_grad, hipStream_t s) {
    net->step++;
    const float b1c = (float)(1.0 - std::pow(0.9, net->step)), b2c = (float)(1.0 - std::pow(0.999, net->step));
    size_t off = 0;
    std::vector<float *> tmp;
    for (auto &Ly : net->L) {
        const size_t np = (size_t)Ly.N * Ly.Kp;
        float *gin = nullptr;
        if (host_grad) {
            std::vector<float> d;
            to_device_layout(Ly, host_grad + off, d);
            gin = dal<float>(np);
            if (!gin) {
                set_error("mlp: device allocation failed");
                return GFSDC_ENOMEM;
            }
            tmp.push_back(gin);
            GFSDC_CHECK(hipMemcpyAsync(gin, d.data(), sizeof(float) * np, hipMemcpyHostToDevice, s));
            GFSDC_CHECK(hipStreamSynchronize(s));
        }
        hipLaunchKernelGGL(k_adamw, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, s, Ly.w, Ly.m, Ly.v, Ly.gpart,
                           net->last_S, (const float *)gin, Ly.N, Ly.K, Ly.Kp, (float)lr, (float)wd, b1c, b2c, Ly.wt,
                           Ly.wk);
        off += (size_t)Ly.K * Ly.Nr + Ly.Nr;
    }
    GFSDC_CHECK(hipGetLastError());
    if (!tmp.empty()) {
        GFSDC_CHECK(hipStreamSynchronize(s));
        for (float *p : tmp) (void)hipFree(p);
    }
    return 0;
}

}  // namespace gfsdc
