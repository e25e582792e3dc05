// Spherical-harmonic transform kernels for gfx950 (CDNA4).
//
// Synthesis of a batch of fields (spectral -> grid) runs as
//   gather      spectral state -> extended spectral columns (n <= T+1), applying the
//               U = u cos(lat), V = v cos(lat) recurrences for vector fields
//   Legendre    per zonal wavenumber m an MFMA GEMM over (latitude pairs) x (degrees)
//               x (2 * fields), split by the parity of n - m so one GEMM per parity
//               covers both hemispheres (P_n^m(-mu) = (-1)^(n-m) P_n^m(mu))
//   FFT         per latitude pair a complex Stockham FFT in LDS packing the two
//               hemispheres' real rows as z = x_north + i x_south
// and analysis is the mirror image (FFT -> Legendre GEMM -> scatter).
// The arithmetic follows the transforms src/shtns.f90 requests from SHTns:
// orthonormal harmonics without Condon-Shortley phase on a Gauss grid.
#include <climits>
#include <cmath>
#include <cstdlib>

#include "gfsdc_internal.h"

namespace gfsdc {

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef double doublex4 __attribute__((ext_vector_type(4)));

template <typename T> struct MF;
// Both take A[i][k] from lane i + 16k and B[k][j] from lane j + 16k; the row of
// accumulator element r of a lane differs: 4*(lane/16) + r for the f32 form,
// lane/16 + 4*r for the f64 form (measured on gfx950).
template <> struct MF<float> {
    typedef floatx4 acc;
    __device__ static inline acc mma(float a, float b, acc c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    __device__ static inline int row(int lane, int r) { return (lane >> 4) * 4 + r; }
};
template <> struct MF<double> {
    typedef doublex4 acc;
    __device__ static inline acc mma(double a, double b, acc c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    __device__ static inline int row(int lane, int r) { return (lane >> 4) + 4 * r; }
};

// ---------------------------------------------------------------------------
// Table generation (FP64 arithmetic on the device)

// Gauss-Legendre nodes mu (cos colatitude, north -> south) and weights (sum 2)
// by Newton iteration on P_nlat(cos theta) in the colatitude, one thread per
// root pair; w = 2 sin^2(theta) / (n (P_{n-1} - x P_n))^2 keeps polar weights exact.
__global__ void k_gauss(double *mu, double *w, int nlat) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (nlat + 1) / 2) return;
    const int n = nlat;
    double th = M_PI * (i + 0.75) / (n + 0.5);
    double x, p0, p1;
    for (int it = 0; it < 100; it++) {
        x = cos(th);
        p0 = 1.0;
        p1 = x;
        for (int k = 2; k <= n; k++) {
            double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
            p0 = p1;
            p1 = p2;
        }
        double dth = -p1 * sin(th) / (n * (p0 - x * p1));
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

__device__ inline double eps_nm_d(int n, int m) {
    if (n <= m) return 0.0;
    return sqrt(((double)n * n - (double)m * m) / (4.0 * n * n - 1.0));
}

// Pbar_n^m(mu_j), n = m..T+1, FP64 three-term recurrence in extended exponent
// (2^500 steps) so high orders near the poles neither underflow early nor
// lose the values that grow back; stored in the parity-split table layout.
template <typename T>
__global__ void k_legendre_table(T *plm, const long *poff, const double *mu, int nhalf, int nhp,
                                 int ntrunc) {
    int m = blockIdx.y;
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nhp) return;
    int K = ntrunc + 2 - m, Ke = (K + 1) >> 1;
    T *pe = plm + poff[m], *po = pe + (long)Ke * nhp;
    if (j >= nhalf) {
        for (int i = 0; i < K; i++) ((i & 1) ? po : pe)[(long)(i >> 1) * nhp + j] = T(0);
        return;
    }
    const double BIG = 0x1p500, SMALL = 0x1p-500;
    double x = mu[j];
    double s = sqrt((1.0 - x) * (1.0 + x));
    double p = sqrt(1.0 / (4.0 * M_PI));
    int e = 0;
    for (int k = 1; k <= m; k++) {
        p *= sqrt((2.0 * k + 1.0) / (2.0 * k)) * s;
        if (p != 0.0 && fabs(p) < SMALL) {
            p *= BIG;
            e--;
        }
    }
    double pm2 = 0.0, pm1 = p;
    for (int i = 0; i < K; i++) {
        int n = m + i;
        double val;
        if (i == 0) {
            val = pm1;
        } else {
            val = (x * pm1 - eps_nm_d(n - 1, m) * pm2) / eps_nm_d(n, m);
            pm2 = pm1;
            pm1 = val;
            if (e < 0 && fabs(val) > BIG) {
                pm1 *= SMALL;
                pm2 *= SMALL;
                val *= SMALL;
                e++;
            }
        }
        double out = (e == 0) ? val : (e == -1 ? val * SMALL : 0.0);
        ((i & 1) ? po : pe)[(long)(i >> 1) * nhp + j] = (T)out;
    }
}

// ---------------------------------------------------------------------------
// Legendre GEMMs. One workgroup = one order m, a 64-row x 128-column output tile;
// 4 waves in a 2x2 arrangement, each wave 32 x 64 via 2 x 4 MFMA 16x16x4 blocks.
// The synthesis gives each parity of n - m its own workgroup (half the
// accumulators: 7 waves per SIMD instead of 3); the analysis keeps both
// parities in one workgroup, which measured faster there. Global loads are 16-byte vectors,
// prefetched into registers for chunk k+1 while chunk k is computed from LDS.
// FP32 swaps the MFMA operands (C^T = B^T A^T) so the four accumulators of a lane
// are four consecutive columns of one output row and the epilogue stores 16-byte
// vectors; FP64 keeps rows in the accumulator and consecutive lanes store
// consecutive columns.
constexpr int LDS_PAD = 16;

template <typename T> struct Vec;
template <> struct Vec<float> {
    typedef float4 t;
    static constexpr int n = 4;
    static constexpr bool swap = true;
};
template <> struct Vec<double> {
    typedef double2 t;
    static constexpr int n = 2;
    static constexpr bool swap = false;
};

// output columns per workgroup: 128 for FP32, 64 for FP64 (whose accumulators
// take twice the registers); each wave covers half of them
template <typename T> struct GT { static constexpr int BN = GEMM_BN, WB = 4; };
template <> struct GT<double> { static constexpr int BN = 64, WB = 2; };

template <typename T>
__device__ inline void vstore_lds(T *dst, const typename Vec<T>::t &v) {
    *(typename Vec<T>::t *)dst = v;
}
// vector component q (a constant after unrolling) without taking the vector's address
__device__ inline float vget(const float4 &v, int q) { return q == 0 ? v.x : q == 1 ? v.y : q == 2 ? v.z : v.w; }
__device__ inline double vget(const double2 &v, int q) { return q == 0 ? v.x : v.y; }
// x, or zero when keep is false, selected component by component (a select of
// whole vector structs is lowered to a select of addresses through scratch)
__device__ inline float4 vkeep(const float4 &x, bool keep) {
    return float4{keep ? x.x : 0.f, keep ? x.y : 0.f, keep ? x.z : 0.f, keep ? x.w : 0.f};
}
__device__ inline double2 vkeep(const double2 &x, bool keep) { return double2{keep ? x.x : 0.0, keep ? x.y : 0.0}; }

// One LDS chunk of a wave's 32 x (16 WB) tile. GUARD skips (wave-uniformly) the
// k-steps of 4 at or beyond kmax and the 16-column blocks at or beyond nbj, so a
// partial last chunk or column tile issues only the MFMAs it needs.
template <typename T, bool GUARD, typename Acc>
__device__ inline void gemm_chunk(T (*As)[GEMM_BM + LDS_PAD], T (*Bs)[GT<T>::BN + LDS_PAD], Acc (&E)[2][GT<T>::WB],
                                  int lane, int wm, int wn, int kmax = GEMM_BK, int nbj = GT<T>::WB) {
    constexpr int NK = GEMM_BK / 4, WB = GT<T>::WB;
    const int li = lane & 15, kl = lane >> 4;
    if constexpr (!GUARD) {
        // all operands of the chunk first, so the LDS reads overlap each other
        // and the matrix-core work instead of waiting before every k-step
        T a[NK][2], b[NK][WB];
#pragma unroll
        for (int q = 0; q < NK; q++) {
#pragma unroll
            for (int bi = 0; bi < 2; bi++) a[q][bi] = As[4 * q + kl][wm + bi * 16 + li];
#pragma unroll
            for (int bj = 0; bj < WB; bj++) b[q][bj] = Bs[4 * q + kl][wn + bj * 16 + li];
        }
#pragma unroll
        for (int q = 0; q < NK; q++)
#pragma unroll
            for (int bj = 0; bj < WB; bj++)
#pragma unroll
                for (int bi = 0; bi < 2; bi++) {
                    if (Vec<T>::swap)
                        E[bi][bj] = MF<T>::mma(b[q][bj], a[q][bi], E[bi][bj]);
                    else
                        E[bi][bj] = MF<T>::mma(a[q][bi], b[q][bj], E[bi][bj]);
                }
        return;
    }
#pragma unroll
    for (int ks = 0; ks < GEMM_BK; ks += 4) {
        if (GUARD && ks >= kmax) break;
        const int kr = ks + kl;
        T a[2];
#pragma unroll
        for (int bi = 0; bi < 2; bi++) a[bi] = As[kr][wm + bi * 16 + li];
#pragma unroll
        for (int bj = 0; bj < WB; bj++) {
            if (GUARD && bj >= nbj) break;
            T b = Bs[kr][wn + bj * 16 + li];
#pragma unroll
            for (int bi = 0; bi < 2; bi++) {
                if (Vec<T>::swap)
                    E[bi][bj] = MF<T>::mma(b, a[bi], E[bi][bj]);
                else
                    E[bi][bj] = MF<T>::mma(a[bi], b, E[bi][bj]);
            }
        }
    }
}

// Synthesis, one parity per workgroup: the even-degree sum E_j goes to row j of
// fs[m], the odd-degree sum O_j to row nlat-1-j; the Fourier synthesis forms
// the hemispheres E + O (north) and E - O (south) as it loads them. Each
// workgroup holds one parity's accumulators only. In FP32 the output tile is
// restaged through LDS after the k-loop so that each store instruction writes
// whole 128-byte lines of the Fourier buffer (8 lines per instruction instead
// of 16 half lines): the stores dominated the memory pipeline of this kernel.
template <typename T>
__global__ __launch_bounds__(256) void k_leg_synth(const T *__restrict__ plm, const long *__restrict__ poff,
                                                   const int *__restrict__ keoff, const T *__restrict__ sx,
                                                   int ldx, T *__restrict__ fs, int ldf, int nhalf, int nhp,
                                                   int nlat, int ntrunc, int ntj, int ncol,
                                                   const int *__restrict__ nfdyn) {
    typedef typename MF<T>::acc acc_t;
    typedef typename Vec<T>::t V;
    constexpr int VW = Vec<T>::n;
    constexpr int BN = GT<T>::BN, WB = GT<T>::WB;
    constexpr int NA = GEMM_BK * GEMM_BM / VW / 256, NB = GEMM_BK * BN / VW / 256;
    constexpr bool STG = Vec<T>::swap;
    // logical tile ((i * 2 + par) * ntj + tj) * ntc + tc, contiguous per XCD: the
    // column and latitude tiles of one (m, parity) share its Legendre block and
    // spectral columns in that XCD's L2. Orders are taken as m = 0, T, 1, T-1, ...
    // so that every XCD's range holds a like amount of work (K = T + 2 - m).
    const int ntc = gridDim.x / (2 * (ntrunc + 1) * ntj);
    const int L = xcd_tile(blockIdx.x, gridDim.x);
    const int tc = L % ntc, g = L / ntc, tj = g % ntj, ip = g / ntj, par = ip & 1, i = ip >> 1;
    const int m = (i & 1) ? ntrunc - (i >> 1) : (i >> 1);
    const int K = ntrunc + 2 - m, Ke = (K + 1) >> 1, Ko = K >> 1, Kp = par ? Ko : Ke;
    const T *P = plm + poff[m] + (par ? (long)Ke * nhp : 0L);
    const T *S = sx + (long)keoff[m] * ldx + (long)par * ldx;
    const int j0 = tj * GEMM_BM, c0 = tc * BN;
    if (nfdyn) ncol = min(ncol, 2 * *nfdyn);
    if (c0 >= ncol) return;  // a column tile past the device field count
    // operand chunks; after the k-loop (STG) each wave's 16 x 64 output slab
    constexpr int STG_P = BN / 2 + 4;
    constexpr int LDS_OPS = GEMM_BK * (GEMM_BM + LDS_PAD) + GEMM_BK * (BN + LDS_PAD);
    constexpr int LDS_STG = STG ? 4 * 16 * STG_P : 0;
    __shared__ __attribute__((aligned(16))) T lds[LDS_OPS > LDS_STG ? LDS_OPS : LDS_STG];
    T (*As)[GEMM_BM + LDS_PAD] = (T (*)[GEMM_BM + LDS_PAD])lds;
    T (*Bs)[BN + LDS_PAD] = (T (*)[BN + LDS_PAD])(lds + GEMM_BK * (GEMM_BM + LDS_PAD));
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * (BN / 2);
    acc_t E[2][WB];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < WB; b++) E[a][b] = acc_t{0, 0, 0, 0};
    V ra[NA], rb[NB];
    // clamped in-bounds loads; the rows past Kp are zeroed by value when stored
    // (no select between pointers, and no wait on the prefetch before the MFMAs)
    auto load = [&](int k0) {
#pragma unroll
        for (int u = 0; u < NA; u++) {
            int v = tid + 256 * u, kk = v / (GEMM_BM / VW), jj = (v % (GEMM_BM / VW)) * VW, r = k0 + kk;
            ra[u] = *(const V *)(P + (long)min(r, Kp - 1) * nhp + j0 + jj);
        }
#pragma unroll
        for (int u = 0; u < NB; u++) {
            int v = tid + 256 * u, kk = v / (BN / VW), cc = (v % (BN / VW)) * VW, r = k0 + kk;
            rb[u] = *(const V *)(S + (long)(2 * min(r, Kp - 1)) * ldx + c0 + cc);
        }
    };
    auto store = [&](int k0) {
#pragma unroll
        for (int u = 0; u < NA; u++) {
            int v = tid + 256 * u, kk = v / (GEMM_BM / VW), jj = (v % (GEMM_BM / VW)) * VW;
            vstore_lds<T>(&As[kk][jj], vkeep(ra[u], k0 + kk < Kp));
        }
#pragma unroll
        for (int u = 0; u < NB; u++) {
            int v = tid + 256 * u, kk = v / (BN / VW), cc = (v % (BN / VW)) * VW;
            vstore_lds<T>(&Bs[kk][cc], vkeep(rb[u], k0 + kk < Kp));
        }
    };
    // FP32 prefetches chunk k+1 during chunk k; FP64 (twice the registers per
    // element) loads each chunk just before storing it to keep occupancy
    constexpr bool PREFETCH = sizeof(T) == 4;
    // valid 16-column blocks of this wave (the last column tile is partial)
    const int nbj = min(WB, max(0, (ncol - c0 - wn + 15) / 16));
    if (PREFETCH && Kp > 0) load(0);
    for (int k0 = 0; k0 < Kp; k0 += GEMM_BK) {
        if (!PREFETCH) load(k0);
        store(k0);
        __syncthreads();
        if (PREFETCH && k0 + GEMM_BK < Kp) load(k0 + GEMM_BK);
        if (nbj == WB && k0 + GEMM_BK <= Kp)
            gemm_chunk<T, false>(As, Bs, E, lane, wm, wn);
        else
            gemm_chunk<T, true>(As, Bs, E, lane, wm, wn, Kp - k0, nbj);
        __syncthreads();
    }
    const int nm1 = ntrunc + 1;
    if constexpr (STG) {
        // per 16-row slab bi: lanes hold row lane % 16, columns 4 (lane / 16) + r of each
        // 16 x 16 block; restaged, lane holds columns 4 (lane % 16) + r of row 4 q + lane / 16,
        // so 16 lanes write 64 consecutive columns (two whole lines) of one row. The slab
        // area of each wave is its own; LDS operations of one wave complete in order.
        T *st = lds + wave * 16 * STG_P;
        const int c = c0 + wn + 4 * (lane & 15);
#pragma unroll
        for (int bi = 0; bi < 2; bi++) {
#pragma unroll
            for (int bj = 0; bj < WB; bj++) {
                acc_t e = E[bi][bj];
                *(float4 *)(st + (lane & 15) * STG_P + bj * 16 + 4 * (lane >> 4)) = float4{e[0], e[1], e[2], e[3]};
            }
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int r = 4 * q + (lane >> 4), j = j0 + wm + bi * 16 + r;
                const float4 v = *(const float4 *)(st + r * STG_P + 4 * (lane & 15));
                if (j < nhalf && c < ncol) *(float4 *)(fs + fs_index(par ? nlat - 1 - j : j, m, c, ldf, nm1)) = v;
            }
        }
    } else {
#pragma unroll
        for (int bi = 0; bi < 2; bi++)
#pragma unroll
            for (int bj = 0; bj < WB; bj++)
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    int j = j0 + wm + bi * 16 + MF<T>::row(lane, r);
                    int c = c0 + wn + bj * 16 + (lane & 15);
                    if (j < nhalf && c < ncol) fs[fs_index(par ? nlat - 1 - j : j, m, c, ldf, nm1)] = E[bi][bj][r];
                }
    }
}

// One LDS chunk of one parity (analysis). GUARD skips, wave-uniformly, the
// 16-row blocks at or beyond ne (rows past the parity's last degree) and the
// 16-column blocks at or beyond nbj (columns past the last field). A is
// k-major, As[k * (GEMM_BM + LDS_PAD) + i]; B is Bs[k * bp + c].
template <typename T, bool GUARD, int WB, typename Acc>
__device__ inline void gemm_chunk1(const T *As, const T *Bs, int bp, Acc (&E)[2][WB], int lane, int wm, int wn,
                                   int ne = 2, int nbj = WB) {
#pragma unroll
    for (int ks = 0; ks < GEMM_BK; ks += 4) {
        const int kr = ks + (lane >> 4), li = lane & 15;
        T a[2];
#pragma unroll
        for (int bi = 0; bi < 2; bi++) a[bi] = As[kr * (GEMM_BM + LDS_PAD) + wm + bi * 16 + li];
#pragma unroll
        for (int bj = 0; bj < WB; bj++) {
            if (GUARD && bj >= nbj) break;
            T b = Bs[kr * bp + wn + bj * 16 + li];
#pragma unroll
            for (int bi = 0; bi < 2; bi++) {
                if (Vec<T>::swap) {
                    if (!GUARD || bi < ne) E[bi][bj] = MF<T>::mma(b, a[bi], E[bi][bj]);
                } else {
                    if (!GUARD || bi < ne) E[bi][bj] = MF<T>::mma(a[bi], b, E[bi][bj]);
                }
            }
        }
    }
}

__host__ __device__ inline int round_up_d(int a, int b) { return (a + b - 1) / b * b; }

// plmt[tpoff[m] + (par nhp + j) Kw + i'] = P(m, n = m + 2 i' + par)(lat j), with
// Kw = roundup(Ke, GEMM_BM); zero for i' past the parity's last degree.
// Grid (j, par, m), GEMM_BM threads striding over i'.
template <typename T>
__global__ void k_transpose_plm(const T *__restrict__ plm, const long *__restrict__ poff, T *__restrict__ plmt,
                                const long *__restrict__ tpoff, int nhp, int ntrunc) {
    const int j = blockIdx.x, par = blockIdx.y, m = blockIdx.z;
    const int K = ntrunc + 2 - m, Ke = (K + 1) >> 1, Kp = par ? (K >> 1) : Ke, Kw = round_up_d(Ke, GEMM_BM);
    const T *P = plm + poff[m] + (par ? (long)Ke * nhp : 0L);
    T *D = plmt + tpoff[m] + ((long)par * nhp + j) * Kw;
    for (int i = threadIdx.x; i < Kw; i += blockDim.x) D[i] = i < Kp ? P[(long)i * nhp + j] : T(0);
}

// Analysis, one parity per workgroup: output degrees i' of that parity, K over
// the (zero-padded) latitude pairs. fa[m][0][j][c] = w_j (F_N + F_S),
// fa[m][1][j][c] = w_j (F_N - F_S). Each workgroup holds one parity's
// accumulators only (half the registers and LDS of a two-parity tile). The
// Legendre operand comes from the latitude-major copy of the table (plmt), so
// one chunk is GEMM_BK rows of GEMM_BM consecutive degrees: whole-line loads
// and 16-byte LDS stores (the degree-major table gave 64-byte pieces).
template <typename T>
__global__ __launch_bounds__(256) void k_leg_anal(const T *__restrict__ plmt, const long *__restrict__ tpoff,
                                                  const int *__restrict__ keoff, const int *__restrict__ mtile_m,
                                                  const int *__restrict__ mtile_i, const T *__restrict__ fa,
                                                  int ldf, T *__restrict__ sx, int ldx, int nhp, int ntrunc, int ncol,
                                                  int nmt, const int *__restrict__ nfdyn) {
    typedef typename MF<T>::acc acc_t;
    typedef typename Vec<T>::t V;
    constexpr int VW = Vec<T>::n;
    constexpr int BN = GT<T>::BN, WB = GT<T>::WB;
    constexpr int NA = GEMM_BK * GEMM_BM / VW / 256, NB = GEMM_BK * BN / VW / 256;
    // logical tile (mt * 2 + parity) * ntc + tc, contiguous per XCD: the column
    // tiles of one (m, degree tile, parity) share its Legendre block in that XCD's L2
    const int ntc = gridDim.x / (2 * nmt);
    const int L = xcd_tile(blockIdx.x, gridDim.x), tc = L % ntc, g = L / ntc, par = g & 1, mt = g >> 1;
    const int m = mtile_m[mt], i0 = mtile_i[mt];
    const int K = ntrunc + 2 - m, Ke = (K + 1) >> 1, Ko = K >> 1, Kp = par ? Ko : Ke;
    if (i0 >= Kp) return;  // an odd-parity tile past the last odd degree (the tiles follow the even block)
    const T *F = fa + (long)(2 * m + par) * nhp * ldf;
    const int c0 = tc * BN;
    if (nfdyn) ncol = min(ncol, 2 * *nfdyn);
    if (c0 >= ncol) return;  // a column tile past the device field count
    // [parity][j][i'] rows of Kw = roundup(Ke, GEMM_BM) degrees, zero past the
    // parity's last degree
    const int Kw = round_up_d(Ke, GEMM_BM);
    const T *PT = plmt + tpoff[m] + (long)par * nhp * Kw + i0;
    __shared__ __attribute__((aligned(16))) T As[GEMM_BK][GEMM_BM + LDS_PAD];
    __shared__ __attribute__((aligned(16))) T Bs[GEMM_BK][BN + LDS_PAD];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * (BN / 2);
    acc_t E[2][WB];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < WB; b++) E[a][b] = acc_t{0, 0, 0, 0};
    // vectors are assembled component by component (vkeep): a plain copy of the
    // vector struct from memory puts these arrays in scratch
    V ra[NA], rb[NB];
    auto load = [&](int j0) {
#pragma unroll
        for (int u = 0; u < NA; u++) {
            int v = tid + 256 * u, kk = v / (GEMM_BM / VW), ii = (v % (GEMM_BM / VW)) * VW;
            ra[u] = vkeep(*(const V *)(PT + (long)(j0 + kk) * Kw + ii), true);
        }
#pragma unroll
        for (int u = 0; u < NB; u++) {
            int v = tid + 256 * u, kk = v / (BN / VW), cc = (v % (BN / VW)) * VW;
            rb[u] = vkeep(*(const V *)(F + (long)(j0 + kk) * ldf + c0 + cc), true);
        }
    };
    auto store = [&]() {
#pragma unroll
        for (int u = 0; u < NA; u++) {
            int v = tid + 256 * u, kk = v / (GEMM_BM / VW), ii = (v % (GEMM_BM / VW)) * VW;
            vstore_lds<T>(&As[kk][ii], ra[u]);
        }
#pragma unroll
        for (int u = 0; u < NB; u++) {
            int v = tid + 256 * u, kk = v / (BN / VW), cc = (v % (BN / VW)) * VW;
            vstore_lds<T>(&Bs[kk][cc], rb[u]);
        }
    };
    // FP32 prefetches chunk k+1 during chunk k; FP64 (twice the registers per
    // element) loads each chunk just before storing it to keep occupancy
    constexpr bool PREFETCH = sizeof(T) == 4;
    // valid 16-row blocks and 16-column blocks of this wave
    const int ne = min(2, max(0, (Kp - i0 - wm + 15) / 16));
    const int nbj = min(WB, max(0, (ncol - c0 - wn + 15) / 16));
    const bool full = ne == 2 && nbj == WB;
    if (PREFETCH) load(0);
    for (int j0 = 0; j0 < nhp; j0 += GEMM_BK) {
        if (!PREFETCH) load(j0);
        store();
        __syncthreads();
        if (PREFETCH && j0 + GEMM_BK < nhp) load(j0 + GEMM_BK);
        if (full)
            gemm_chunk1<T, false, WB>(&As[0][0], Bs[0], BN + LDS_PAD, E, lane, wm, wn);
        else
            gemm_chunk1<T, true, WB>(&As[0][0], Bs[0], BN + LDS_PAD, E, lane, wm, wn, ne, nbj);
        __syncthreads();
    }
    T *X = sx + (long)keoff[m] * ldx;
#pragma unroll
    for (int bi = 0; bi < 2; bi++)
#pragma unroll
        for (int bj = 0; bj < WB; bj++) {
            if constexpr (Vec<T>::swap) {
                int i = i0 + wm + bi * 16 + (lane & 15);
                int c = c0 + wn + bj * 16 + 4 * (lane >> 4);
                acc_t e = E[bi][bj];
                if (i < Kp) *(float4 *)(X + (long)(2 * i + par) * ldx + c) = float4{e[0], e[1], e[2], e[3]};
            } else {
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    int i = i0 + wm + bi * 16 + MF<T>::row(lane, r);
                    int c = c0 + wn + bj * 16 + (lane & 15);
                    if (i < Kp) X[(long)(2 * i + par) * ldx + c] = E[bi][bj][r];
                }
            }
        }
}

// ---------------------------------------------------------------------------
// BF16x3 Legendre GEMMs (FP32 plan option): each FP32 operand x is split into
// bf16 hi = rn(x) and lo = rn(x - hi); a product is hi.hi + hi.lo + lo.hi
// (the lo.lo term, ~2^-16 relative, is dropped) on the bf16 matrix cores
// (v_mfma_f32_16x16x32_bf16, FP32 accumulation). Same tiling as the FP32
// kernels (64 x 128 per workgroup, both parities); operands are staged in LDS
// k-contiguous ([row][k], 8 consecutive k per lane and MFMA) and the MFMA
// operands are swapped so a lane's four accumulators are four consecutive
// output columns (16-byte stores).
typedef unsigned short b3u8 __attribute__((ext_vector_type(8)));
typedef __bf16 b3bf8 __attribute__((ext_vector_type(8)));
typedef float b3f4 __attribute__((ext_vector_type(4)));
constexpr int B3K = 32, B3P = B3K + 8;

__device__ inline unsigned short b3_rn(float f) {
    unsigned int u = __builtin_bit_cast(unsigned int, f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__device__ inline void b3_split(float x, unsigned short &hi, unsigned short &lo) {
    hi = b3_rn(x);
    lo = b3_rn(x - __builtin_bit_cast(float, (unsigned int)hi << 16));
}
__device__ inline b3f4 b3_mma3(const unsigned short *bh, const unsigned short *bl, const unsigned short *ah,
                               const unsigned short *al, b3f4 c) {
    const b3bf8 Bh = __builtin_bit_cast(b3bf8, *(const b3u8 *)bh), Bl = __builtin_bit_cast(b3bf8, *(const b3u8 *)bl);
    const b3bf8 Ah = __builtin_bit_cast(b3bf8, *(const b3u8 *)ah), Al = __builtin_bit_cast(b3bf8, *(const b3u8 *)al);
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Bl, Ah, c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Bh, Al, c, 0, 0, 0);
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(Bh, Ah, c, 0, 0, 0);
}

__global__ __launch_bounds__(256) void k_leg_synth_b3(const float *__restrict__ plm, const long *__restrict__ poff,
                                                      const int *__restrict__ keoff, const float *__restrict__ sx,
                                                      int ldx, float *__restrict__ fs, int ldf, int nhalf, int nhp,
                                                      int nlat, int ntrunc, int ntj) {
    const int tc = blockIdx.x, mt = blockIdx.y;
    const int m = mt / ntj, tj = mt - m * ntj;
    const int K = ntrunc + 2 - m, Ke = (K + 1) >> 1, Ko = K >> 1;
    const float *P0 = plm + poff[m], *P1 = P0 + (long)Ke * nhp;
    const float *S = sx + (long)keoff[m] * ldx;
    const int j0 = tj * GEMM_BM, c0 = tc * GEMM_BN;
    __shared__ __attribute__((aligned(16))) unsigned short Ah[2][GEMM_BM][B3P], Al[2][GEMM_BM][B3P];
    __shared__ __attribute__((aligned(16))) unsigned short Bh[2][GEMM_BN][B3P], Bl[2][GEMM_BN][B3P];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 64;
    b3f4 E[2][4], O[2][4];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 4; b++) E[a][b] = O[a][b] = b3f4{0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < Ke; k0 += B3K) {
#pragma unroll
        for (int par = 0; par < 2; par++) {
            const float *Pp = par ? P1 : P0;
            const int Kpp = par ? Ko : Ke;
#pragma unroll
            for (int u = 0; u < 2; u++) {  // P: 32 degrees x 64 latitudes
                const int v = tid + 256 * u, kk = v >> 4, jj = (v & 15) * 4, i = k0 + kk;
                float4 x = *(const float4 *)(Pp + (long)min(i, Kpp - 1) * nhp + j0 + jj);
                x = vkeep(x, i < Kpp);
                b3_split(x.x, Ah[par][jj][kk], Al[par][jj][kk]);
                b3_split(x.y, Ah[par][jj + 1][kk], Al[par][jj + 1][kk]);
                b3_split(x.z, Ah[par][jj + 2][kk], Al[par][jj + 2][kk]);
                b3_split(x.w, Ah[par][jj + 3][kk], Al[par][jj + 3][kk]);
            }
#pragma unroll
            for (int u = 0; u < 4; u++) {  // S: 32 degrees x 128 columns
                const int v = tid + 256 * u, kk = v >> 5, cc = (v & 31) * 4, i = k0 + kk;
                float4 x = *(const float4 *)(S + (long)(2 * min(i, Kpp - 1) + par) * ldx + c0 + cc);
                x = vkeep(x, i < Kpp);
                b3_split(x.x, Bh[par][cc][kk], Bl[par][cc][kk]);
                b3_split(x.y, Bh[par][cc + 1][kk], Bl[par][cc + 1][kk]);
                b3_split(x.z, Bh[par][cc + 2][kk], Bl[par][cc + 2][kk]);
                b3_split(x.w, Bh[par][cc + 3][kk], Bl[par][cc + 3][kk]);
            }
        }
        __syncthreads();
        const int r = lane & 15, kq = 8 * (lane >> 4);
#pragma unroll
        for (int bi = 0; bi < 2; bi++)
#pragma unroll
            for (int bj = 0; bj < 4; bj++) {
                const int jr = wm + bi * 16 + r, cr = wn + bj * 16 + r;
                E[bi][bj] = b3_mma3(&Bh[0][cr][kq], &Bl[0][cr][kq], &Ah[0][jr][kq], &Al[0][jr][kq], E[bi][bj]);
                O[bi][bj] = b3_mma3(&Bh[1][cr][kq], &Bl[1][cr][kq], &Ah[1][jr][kq], &Al[1][jr][kq], O[bi][bj]);
            }
        __syncthreads();
    }
#pragma unroll
    for (int bi = 0; bi < 2; bi++)
#pragma unroll
        for (int bj = 0; bj < 4; bj++) {
            const int j = j0 + wm + bi * 16 + (lane & 15), c = c0 + wn + bj * 16 + 4 * (lane >> 4);
            if (j < nhalf) {  // even sum to row j, odd sum to row nlat-1-j (as k_leg_synth)
                const b3f4 e = E[bi][bj], o = O[bi][bj];
                *(float4 *)(fs + fs_index(j, m, c, ldf, ntrunc + 1)) = float4{e[0], e[1], e[2], e[3]};
                *(float4 *)(fs + fs_index(nlat - 1 - j, m, c, ldf, ntrunc + 1)) = float4{o[0], o[1], o[2], o[3]};
            }
        }
}

__global__ __launch_bounds__(256) void k_leg_anal_b3(const float *__restrict__ plm, const long *__restrict__ poff,
                                                     const int *__restrict__ keoff, const int *__restrict__ mtile_m,
                                                     const int *__restrict__ mtile_i, const float *__restrict__ fa,
                                                     int ldf, float *__restrict__ sx, int ldx, int nhp, int ntrunc) {
    const int tc = blockIdx.x;
    const int m = mtile_m[blockIdx.y], i0 = mtile_i[blockIdx.y];
    const int K = ntrunc + 2 - m, Ke = (K + 1) >> 1, Ko = K >> 1;
    const float *P0 = plm + poff[m], *P1 = P0 + (long)Ke * nhp;
    const float *F0 = fa + (long)(2 * m) * nhp * ldf, *F1 = F0 + (long)nhp * ldf;
    const int c0 = tc * GEMM_BN;
    __shared__ __attribute__((aligned(16))) unsigned short Ah[2][GEMM_BM][B3P], Al[2][GEMM_BM][B3P];
    __shared__ __attribute__((aligned(16))) unsigned short Bh[2][GEMM_BN][B3P], Bl[2][GEMM_BN][B3P];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 64;
    b3f4 E[2][4], O[2][4];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 4; b++) E[a][b] = O[a][b] = b3f4{0.f, 0.f, 0.f, 0.f};
    for (int j0 = 0; j0 < nhp; j0 += B3K) {
#pragma unroll
        for (int par = 0; par < 2; par++) {
            const float *Pp = par ? P1 : P0;
            const int Kpp = par ? Ko : Ke;
#pragma unroll
            for (int u = 0; u < 2; u++) {  // P: 64 degrees x 32 latitudes (k-contiguous already)
                const int v = tid + 256 * u, ii = v >> 3, jj = (v & 7) * 4, i = i0 + ii;
                float4 x = *(const float4 *)(Pp + (long)min(i, Kpp - 1) * nhp + j0 + jj);
                x = vkeep(x, i < Kpp);
                b3_split(x.x, Ah[par][ii][jj], Al[par][ii][jj]);
                b3_split(x.y, Ah[par][ii][jj + 1], Al[par][ii][jj + 1]);
                b3_split(x.z, Ah[par][ii][jj + 2], Al[par][ii][jj + 2]);
                b3_split(x.w, Ah[par][ii][jj + 3], Al[par][ii][jj + 3]);
            }
#pragma unroll
            for (int u = 0; u < 4; u++) {  // F: 32 latitudes x 128 columns
                const int v = tid + 256 * u, kk = v >> 5, cc = (v & 31) * 4;
                const float4 x = *(const float4 *)((par ? F1 : F0) + (long)(j0 + kk) * ldf + c0 + cc);
                b3_split(x.x, Bh[par][cc][kk], Bl[par][cc][kk]);
                b3_split(x.y, Bh[par][cc + 1][kk], Bl[par][cc + 1][kk]);
                b3_split(x.z, Bh[par][cc + 2][kk], Bl[par][cc + 2][kk]);
                b3_split(x.w, Bh[par][cc + 3][kk], Bl[par][cc + 3][kk]);
            }
        }
        __syncthreads();
        const int r = lane & 15, kq = 8 * (lane >> 4);
#pragma unroll
        for (int bi = 0; bi < 2; bi++)
#pragma unroll
            for (int bj = 0; bj < 4; bj++) {
                const int ir = wm + bi * 16 + r, cr = wn + bj * 16 + r;
                E[bi][bj] = b3_mma3(&Bh[0][cr][kq], &Bl[0][cr][kq], &Ah[0][ir][kq], &Al[0][ir][kq], E[bi][bj]);
                O[bi][bj] = b3_mma3(&Bh[1][cr][kq], &Bl[1][cr][kq], &Ah[1][ir][kq], &Al[1][ir][kq], O[bi][bj]);
            }
        __syncthreads();
    }
    float *X = sx + (long)keoff[m] * ldx;
#pragma unroll
    for (int bi = 0; bi < 2; bi++)
#pragma unroll
        for (int bj = 0; bj < 4; bj++) {
            const int i = i0 + wm + bi * 16 + (lane & 15), c = c0 + wn + bj * 16 + 4 * (lane >> 4);
            const b3f4 e = E[bi][bj], o = O[bi][bj];
            if (i < Ke) *(float4 *)(X + (long)(2 * i) * ldx + c) = float4{e[0], e[1], e[2], e[3]};
            if (i < Ko) *(float4 *)(X + (long)(2 * i + 1) * ldx + c) = float4{o[0], o[1], o[2], o[3]};
        }
}

// ---------------------------------------------------------------------------
// Stockham FFT in LDS: NFFT transforms of length N (radices 2, 3, 4), in place
// with a barrier between the load and store half of every stage.
template <typename T> __device__ __forceinline__ cplx<T> cmul(cplx<T> a, cplx<T> b) {
    return cplx<T>{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}

template <typename T, int R> __device__ __forceinline__ void small_dft(cplx<T> *v, int sign) {
    if (R == 2) {
        cplx<T> a = v[0], b = v[1];
        v[0] = cplx<T>{a.x + b.x, a.y + b.y};
        v[1] = cplx<T>{a.x - b.x, a.y - b.y};
    } else if (R == 3) {
        const T s3 = T(0.86602540378443864676) * T(sign);
        cplx<T> a0 = v[0], a1 = v[1], a2 = v[2];
        cplx<T> t1{a1.x + a2.x, a1.y + a2.y}, t2{a1.x - a2.x, a1.y - a2.y};
        cplx<T> b{a0.x - T(0.5) * t1.x, a0.y - T(0.5) * t1.y};
        v[0] = cplx<T>{a0.x + t1.x, a0.y + t1.y};
        v[1] = cplx<T>{b.x - s3 * t2.y, b.y + s3 * t2.x};
        v[2] = cplx<T>{b.x + s3 * t2.y, b.y - s3 * t2.x};
    } else if (R == 4) {
        cplx<T> a0 = v[0], a1 = v[1], a2 = v[2], a3 = v[3];
        cplx<T> s02{a0.x + a2.x, a0.y + a2.y}, d02{a0.x - a2.x, a0.y - a2.y};
        cplx<T> s13{a1.x + a3.x, a1.y + a3.y}, d13{a1.x - a3.x, a1.y - a3.y};
        // w * d13 with w = sign * i
        cplx<T> wd{-T(sign) * d13.y, T(sign) * d13.x};
        v[0] = cplx<T>{s02.x + s13.x, s02.y + s13.y};
        v[2] = cplx<T>{s02.x - s13.x, s02.y - s13.y};
        v[1] = cplx<T>{d02.x + wd.x, d02.y + wd.y};
        v[3] = cplx<T>{d02.x - wd.x, d02.y - wd.y};
    }
}

// LDS layout of the transforms: field q's element x sits at q * fft_pitch(N) +
// swz(x), where swz permutes x within its block of 16 complex values by XOR
// with key(x / 16), key(v) = (v ^ 4 v) mod 16 (5 v multiplied without carries).
// The early Stockham stages write with stride R (Ns = 1) and in runs of Ns = 4;
// under this permutation both patterns, and the contiguous reads, put 16
// consecutive lanes on 16 different bank pairs. The field pitch is 8 complex
// values past a multiple of 16, so the lanes that alternate between two fields
// (loading and storing the Fourier coefficients) fall on different banks.
// key is linear under XOR: for x = g + C, with C a multiple of 16 whose bits
// above the fourth miss those of g / 16, swz(x) = (g ^ key(g / 16) ^ key(C / 16)) + C,
// a per-lane value XOR a constant, plus a constant the LDS instruction carries.
__host__ __device__ constexpr int fft_pitch(int N) { return ((N + 15) & ~15) + 8; }
__host__ __device__ constexpr int fft_key(int v) { return (v ^ (v << 2)) & 15; }
__device__ inline int fft_pad(int x) { return x ^ fft_key(x >> 4); }

// Split of the element indices one wavefront's Stockham stage reads
// (x = j + r nbf) and writes (x = (j / Ns) Ns R + j % Ns + r Ns), j = lane + 64 bb,
// into a lane part g and a constant part c, x = g + c.
struct XSplit {
    int g, c;
};
constexpr XSplit split_rd(int lane, int bb, int r, int nbf) { return XSplit{lane, 64 * bb + r * nbf}; }
constexpr XSplit split_wr(int lane, int bb, int r, int R, int Ns) {
    if (64 % Ns == 0) {  // j / Ns = lane / Ns + 64 bb / Ns, j % Ns = lane % Ns
        int g = (lane / Ns) * Ns * R + lane % Ns, c = 64 * R * bb;
        if (Ns % 16 == 0)
            c += r * Ns;
        else
            g += r * Ns;
        return XSplit{g, c};
    }
    return XSplit{lane, 64 * bb + r * Ns};  // valid when every j < Ns (a last stage)
}
constexpr bool split_ok(XSplit s, int x) { return s.g + s.c == x && s.c % 16 == 0 && ((s.g >> 4) & (s.c >> 4)) == 0; }
// whether every read (WR = false) or write (WR = true) of the stage splits so
constexpr bool split_all(int N, int R, int Ns, bool WR) {
    const int nbf = N / R, NB = (nbf + 63) / 64;
    for (int bb = 0; bb < NB; bb++)
        for (int r = 0; r < R; r++)
            for (int lane = 0; lane < 64; lane++) {
                const int j = lane + 64 * bb;
                if (j >= nbf) continue;
                const int x = WR ? (j / Ns) * Ns * R + j % Ns + r * Ns : j + r * nbf;
                if (!split_ok(WR ? split_wr(lane, bb, r, R, Ns) : split_rd(lane, bb, r, nbf), x)) return false;
            }
    return true;
}
__device__ __forceinline__ int swz_split(XSplit s) { return ((s.g ^ fft_key(s.g >> 4)) ^ fft_key(s.c >> 4)) + s.c; }
static_assert(split_all(768, 4, 1, true) && split_all(768, 4, 4, true) && split_all(768, 4, 16, true) &&
                  split_all(768, 4, 64, true) && split_all(768, 3, 256, true) && split_all(768, 4, 1, false) &&
                  split_all(768, 3, 256, false),
              "the 768-point stages take the split LDS indices");

template <typename T, int R, int MAXB>
__device__ __forceinline__ void fft_stage(cplx<T> *buf, int N, int nfft, int Ns, const cplx<T> *__restrict__ tw,
                                 int sign) {
    const int nbf = N / R, total = nbf * nfft, NP = fft_pitch(N);
    cplx<T> v[MAXB][R];
#pragma unroll
    for (int bb = 0; bb < MAXB; bb++) {
        int b = threadIdx.x + bb * 256;
        if (b < total) {
            int q = b / nbf, j = b - q * nbf;
#pragma unroll
            for (int r = 0; r < R; r++) v[bb][r] = buf[q * NP + fft_pad(j + r * nbf)];
        }
    }
    __syncthreads();
#pragma unroll
    for (int bb = 0; bb < MAXB; bb++) {
        int b = threadIdx.x + bb * 256;
        if (b < total) {
            int q = b / nbf, j = b - q * nbf;
            int jm = j % Ns;
#pragma unroll
            for (int r = 1; r < R; r++) {
                cplx<T> t = tw[Ns - 1 + (r - 1) * Ns + jm];  // = w^(jm r tstep), stage-ordered table
                if (sign > 0) t.y = -t.y;
                v[bb][r] = cmul(v[bb][r], t);
            }
            small_dft<T, R>(v[bb], sign);
            int idx = (j / Ns) * Ns * R + jm;
#pragma unroll
            for (int r = 0; r < R; r++) buf[q * NP + fft_pad(idx + r * Ns)] = v[bb][r];
        }
    }
    __syncthreads();
}

template <typename T, int NFFT, int MAXN>
__device__ __forceinline__ void fft_run(cplx<T> *buf, int N, const int *rad, int nrad, const cplx<T> *tw, int sign) {
    int Ns = 1;
    for (int s = 0; s < nrad; s++) {
        int R = rad[s];
        if (R == 4)
            fft_stage<T, 4, (NFFT * MAXN / 4 + 255) / 256>(buf, N, NFFT, Ns, tw, sign);
        else if (R == 3)
            fft_stage<T, 3, (NFFT * MAXN / 3 + 255) / 256>(buf, N, NFFT, Ns, tw, sign);
        else
            fft_stage<T, 2, (NFFT * MAXN / 2 + 255) / 256>(buf, N, NFFT, Ns, tw, sign);
        Ns *= R;
    }
}

struct Radices {
    int n, r[16];
};

// The same radix sequence as factor_nlon, at compile time, for the FFT kernels
// specialised to a transform length (constant divisors, unrolled stages).
struct RadPlan {
    int n, r[16];
};
constexpr RadPlan rad_plan(int n) {
    RadPlan p{0, {}};
    while (n % 4 == 0 && n > 4) {
        p.r[p.n++] = 4;
        n /= 4;
    }
    while (n % 3 == 0) {
        p.r[p.n++] = 3;
        n /= 3;
    }
    while (n % 4 == 0) {
        p.r[p.n++] = 4;
        n /= 4;
    }
    while (n % 2 == 0) {
        p.r[p.n++] = 2;
        n /= 2;
    }
    return p;
}

// Orders the LDS accesses of one wavefront: the compiler moves no LDS access
// across it, and the LDS unit serves one wavefront's instructions in issue
// order, so writes before it are seen by reads after it in every lane.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// One Stockham stage with one field per wavefront (4 fields per workgroup):
// every divisor is a compile-time constant, and where the stage's indices split
// (split_all) an LDS address is a per-lane register plus an immediate offset.
// The stage reads element x through ld(x, p) and hands its outputs to
// st(x, p, v), p being x's permuted LDS index. Only this wavefront touches its
// field, so the stages are ordered by wave_lds_sync instead of workgroup
// barriers. The first stage (Ns = 1) has unit twiddles.
// Twiddles of one Stockham stage in registers: R - 1 per butterfly of this
// lane (none in the first stage, whose twiddles are 1), loaded one stage ahead
// of their use so that the loads overlap a whole stage of LDS work.
template <typename T, int R, int N, int Ns> struct StageTw {
    static constexpr int nbf = N / R, NB = (nbf + 63) / 64, NT = Ns > 1 ? R - 1 : 0;
    cplx<T> t[NB][NT > 0 ? NT : 1];
    __device__ __forceinline__ void load(const cplx<T> *__restrict__ tw) {
        if constexpr (NT > 0) {
            const int lane = threadIdx.x & 63;
#pragma unroll
            for (int bb = 0; bb < NB; bb++) {
                const int jm = min(lane + 64 * bb, nbf - 1) % Ns;  // clamped: lanes past nbf load in bounds
#pragma unroll
                for (int r = 1; r < R; r++) t[bb][r - 1] = tw[Ns - 1 + (r - 1) * Ns + jm];
            }
        }
    }
};
struct NoTw {
    template <typename P> __device__ __forceinline__ void load(const P *) {}
};
// Twiddles of a last stage of 256 butterflies run four consecutive butterflies
// per lane (j = 4 lane + bb), loaded during the stage before it
template <typename T, int R, int NS> struct QuadTw {
    cplx<T> t[4][R - 1];
    __device__ __forceinline__ void load(const cplx<T> *__restrict__ tw) {
        const int lane = threadIdx.x & 63;
#pragma unroll
        for (int bb = 0; bb < 4; bb++)
#pragma unroll
            for (int r = 1; r < R; r++) t[bb][r - 1] = tw[NS - 1 + (r - 1) * NS + 4 * lane + bb];
    }
};

template <typename T, int R, int N, int Ns, typename Ld, typename St, typename TN>
__device__ __forceinline__ void fft_stage_w(const cplx<T> *__restrict__ tw, int sign, Ld ld, St st,
                                            const StageTw<T, R, N, Ns> &tc, TN &tn) {
    constexpr int nbf = N / R, NB = (nbf + 63) / 64;
    constexpr bool SR = split_all(N, R, Ns, false), SW = split_all(N, R, Ns, true);
    const int lane = threadIdx.x & 63;
    cplx<T> v[NB][R];
#pragma unroll
    for (int bb = 0; bb < NB; bb++) {
        const int j = lane + 64 * bb;
        if (nbf % 64 == 0 || j < nbf) {
#pragma unroll
            for (int r = 0; r < R; r++) {
                const int x = j + r * nbf;
                v[bb][r] = ld(x, SR ? swz_split(split_rd(lane, bb, r, nbf)) : fft_pad(x));
            }
        }
    }
    tn.load(tw);      // the next stage's twiddles
    wave_lds_sync();  // in place: every read of the stage before any write
#pragma unroll
    for (int bb = 0; bb < NB; bb++) {
        const int j = lane + 64 * bb;
        if (nbf % 64 == 0 || j < nbf) {
            const int jm = j % Ns;
            if constexpr (Ns > 1) {
#pragma unroll
                for (int r = 1; r < R; r++) {
                    cplx<T> t = tc.t[bb][r - 1];
                    if (sign > 0) t.y = -t.y;
                    v[bb][r] = cmul(v[bb][r], t);
                }
            }
            small_dft<T, R>(v[bb], sign);
            const int idx = (j / Ns) * Ns * R + jm;
#pragma unroll
            for (int r = 0; r < R; r++) {
                const int x = idx + r * Ns;
                st(x, SW ? swz_split(split_wr(lane, bb, r, R, Ns)) : fft_pad(x), v[bb][r]);
            }
        }
    }
    wave_lds_sync();
}

// Stages S.. of the length-N transform of field bq (one wavefront), stage S's
// twiddles tc already loaded. FIRST reads the first stage's inputs through ld
// instead of LDS; LAST hands the last stage's outputs to st instead of LDS (the
// last stage writes x = j + r Ns with j < Ns, consecutive across lanes, so a
// global store from it is coalesced).
template <typename T, int N, int S, int Ns, bool FIRST, bool LAST, int SEND, typename Ld, typename St,
          typename TT = NoTw>
__device__ __forceinline__ void fft_stages_w(cplx<T> *bq, const cplx<T> *tw, int sign, Ld ld, St st,
                                             const StageTw<T, rad_plan(N).r[S], N, Ns> &tc, TT *tail = nullptr) {
    constexpr RadPlan P = rad_plan(N);
    constexpr int SE = SEND < P.n ? SEND : P.n;  // stages S .. SE - 1 run here
    if constexpr (S < SE) {
        constexpr int R = P.r[S];
        auto lds_ld = [&](int, int p) { return bq[p]; };
        auto lds_st = [&](int, int p, const cplx<T> &v) { bq[p] = v; };
        auto run = [&](auto &tn) {
            if constexpr (FIRST && S == 0) {
                if constexpr (LAST && S == P.n - 1)
                    fft_stage_w<T, R, N, Ns>(tw, sign, ld, st, tc, tn);
                else
                    fft_stage_w<T, R, N, Ns>(tw, sign, ld, lds_st, tc, tn);
            } else if constexpr (LAST && S == P.n - 1) {
                fft_stage_w<T, R, N, Ns>(tw, sign, lds_ld, st, tc, tn);
            } else {
                fft_stage_w<T, R, N, Ns>(tw, sign, lds_ld, lds_st, tc, tn);
            }
        };
        if constexpr (S + 1 < SE) {
            StageTw<T, P.r[S + 1], N, Ns * R> tn;
            run(tn);
            fft_stages_w<T, N, S + 1, Ns * R, FIRST, LAST, SEND>(bq, tw, sign, ld, st, tn, tail);
        } else if (tail) {
            run(*tail);  // the twiddles of the stage that follows this call's last
        } else {
            NoTw tn;
            run(tn);
        }
    }
}

template <typename T, int NFFT, int N, int S, int Ns>
__device__ __forceinline__ void fft_stages(cplx<T> *buf, const cplx<T> *tw, int sign) {
    constexpr RadPlan P = rad_plan(N);
    if constexpr (S < P.n) {
        constexpr int R = P.r[S];
        fft_stage<T, R, (NFFT * N / R + 255) / 256>(buf, N, NFFT, Ns, tw, sign);
        fft_stages<T, NFFT, N, S + 1, Ns * R>(buf, tw, sign);
    }
}


template <typename T> struct FftCfg;
template <> struct FftCfg<float> { static constexpr int NFFT = 4, MAXN = 768; };
template <> struct FftCfg<double> { static constexpr int NFFT = 4, MAXN = 768; };
static_assert(fft_pitch(FftCfg<double>::MAXN) * 8 * 16 + FftCfg<double>::MAXN * 16 <= 160 * 1024, "FFT LDS budget");

// Fourier synthesis: one workgroup per latitude pair and NFFT fields. The
// twiddle table is staged in LDS once, the Fourier coefficients of both
// hemispheres are loaded in one batch per thread before any LDS write, and
// only the spectral gap T < k < N-T is zero-filled.
// NC > 0 specialises the kernel to transform length NC (compile-time radix
// stages and divisors, LDS sized to NC); NC = 0 handles any length <= MAXN.
// Workgroup size of the FFT kernels: one wavefront per field on the wavefront
// path (a transform length known at compile time, 4 or 8 fields), else 256.
__host__ __device__ constexpr bool fft_wave(int nfft, int nc) { return nc > 0 && (nfft == 4 || nfft == 8); }
__host__ __device__ constexpr int fft_threads(int nfft, int nc) { return fft_wave(nfft, nc) ? 64 * nfft : 256; }

template <typename T, int NFFT, int MAXN, int NC>
__global__ __launch_bounds__(fft_threads(NFFT, NC)) void k_fft_synth(const T *__restrict__ fs, int ldf, T *__restrict__ grid,
                                                   long gstride, int nf, const unsigned char *__restrict__ rflag,
                                                   const T *__restrict__ mean, const int *__restrict__ slot,
                                                   const T *__restrict__ rcos, const cplx<T> *__restrict__ tw,
                                                   int nlon, int nlat, int ntrunc, Radices rd, int fofs,
                                                   const int *__restrict__ nfdyn, FieldMap fmap, T rdx) {
    constexpr int LN = NC ? NC : MAXN;
    // wavefront path; with 8 fields the twiddles are read from the (cached)
    // global table so that three workgroups fit a CU's LDS
    constexpr bool WAVE = fft_wave(NFFT, NC), GTW = WAVE && NFFT == 8;
    constexpr int NT = fft_threads(NFFT, NC);
    __shared__ cplx<T> buf[NFFT * fft_pitch(LN)];
    __shared__ cplx<T> tws[GTW ? 1 : LN];
    const int nfg = (nf + NFFT - 1) / NFFT, tile = xcd_tile(blockIdx.x, gridDim.x);
    const int jp = tile / nfg, f0 = (tile - jp * nfg) * NFFT;
    if (nfdyn) {
        nf = min(nf, *nfdyn);
        if (f0 >= nf) return;  // a field group past the device field count
    }
    const int j = jp, j2 = nlat - 1 - jp, N = NC ? NC : nlon, NP = fft_pitch(N), tid = threadIdx.x;
    // wavefront path: this wave's field, its 1/cos scaling and added mean, loaded
    // with the coefficients (field clamped into range; a wave past nf stores nothing)
    T w_s = T(1), w_c = T(0);
    long w_g = 0;
    // a last stage of 256 butterflies (the 768-point transform: radix 3, Ns = 256)
    // runs with four consecutive butterflies per lane, so that each lane's outputs
    // of one row are four consecutive grid points (one 16-byte store); its
    // twiddles come from the global table during the stage before it
    constexpr RadPlan PS = rad_plan(NC > 0 ? NC : 2);
    constexpr int RL = PS.r[PS.n > 0 ? PS.n - 1 : 0], NSL = (NC > 0 ? NC : 2) / RL;
    constexpr bool QUADL = WAVE && NSL == 256 && RL <= 4;
    // the Legendre fields of the workgroup's grid fields: where they are
    // consecutive (every group but those at a run's or a member's end) the map
    // is taken once for the group, with workgroup-uniform values
    const int fa = min(f0, nf - 1), fb = min(f0 + NFFT - 1, nf - 1);
    bool dxa;
    const int la = fmap_field(fmap, fa, dxa);
    const bool lin = fmap_linear(fmap, fa, fb);
    auto lfield = [&](int q, bool &dx) {
        if (lin) {
            dx = dxa;
            return la + min(q, fb - fa);
        }
        return fmap_field(fmap, min(f0 + q, nf - 1), dx);
    };
    if constexpr (WAVE) {
        const int fw = min(f0 + (tid >> 6), nf - 1);
        bool dxw;
        const int lw = lfield(tid >> 6, dxw);
        const T rc = rcos[j];  // unconditional, so that no load waits for another
        w_s = rflag[fw] ? rc : T(1);
        w_c = mean ? mean[lw] : T(0);
        if (dxw) w_c = T(0);
        w_g = slot ? slot[fw] : fw;
    }
    constexpr int MAXI = (NFFT * (LN / 2) + NT - 1) / NT;
    const int nitems = (ntrunc + 1) * NFFT;
    // every load of the batch is issued before the first is used: the indices
    // are clamped into the buffer (item nitems - 1, field nf - 1) and the values
    // of the clamped lanes are zeroed afterwards, so no load sits in a branch
    // that would make it wait for the one before
    cplx<T> ev[MAXI], ov[MAXI];
    // a thread's items all belong to field tid % NFFT (NT is a multiple of NFFT);
    // field f reads Legendre field fmap(f) + fofs
    bool dxl;
    const int cf = 2 * (lfield((int)(tid % NFFT), dxl) + fofs);
#pragma unroll
    for (int c = 0; c < MAXI; c++) {
        ev[c] = ov[c] = cplx<T>{T(0), T(0)};
        if (c * NT < nitems) {  // uniform
            const int idx = min(tid + c * NT, nitems - 1);
            const int m = idx / NFFT;
            // rows j / j2 hold the even / odd Legendre sums
            ev[c] = *(const cplx<T> *)(fs + fs_index(j, m, cf, ldf, ntrunc + 1));
            ov[c] = *(const cplx<T> *)(fs + fs_index(j2, m, cf, ldf, ntrunc + 1));
        }
    }
    // hemispheres: north E + O, south E - O; an x-gradient times i m / a
    cplx<T> xn[MAXI], xs[MAXI];
#pragma unroll
    for (int c = 0; c < MAXI; c++) {
        const int idx = tid + c * NT, q = idx % NFFT;
        const bool ok = idx < nitems && f0 + q < nf;
        cplx<T> e = ev[c], o = ov[c];
        if (dxl) {
            const T k = T(idx / NFFT) * rdx;
            e = cplx<T>{-k * e.y, k * e.x};
            o = cplx<T>{-k * o.y, k * o.x};
        }
        xn[c] = ok ? cplx<T>{e.x + o.x, e.y + o.y} : cplx<T>{T(0), T(0)};
        xs[c] = ok ? cplx<T>{e.x - o.x, e.y - o.y} : cplx<T>{T(0), T(0)};
    }
    if constexpr (!GTW)
        for (int t = tid; t < N; t += NT) tws[t] = tw[t];
    const int gap = N - 2 * ntrunc - 1;
#pragma unroll
    for (int q = 0; q < NFFT; q++)
        for (int i = tid; i < gap; i += NT) buf[q * NP + fft_pad(ntrunc + 1 + i)] = cplx<T>{T(0), T(0)};
#pragma unroll
    for (int c = 0; c < MAXI; c++) {
        int idx = tid + c * NT;
        if (idx < nitems) {
            int m = idx / NFFT, q = idx - m * NFFT;
            T xr = xn[c].x, xi = xn[c].y, yr = xs[c].x, yi = xs[c].y;
            if (m == 0) {
                xi = T(0);
                yi = T(0);
            }
            buf[q * NP + fft_pad(m)] = cplx<T>{xr - yi, xi + yr};
            if (m > 0) buf[q * NP + fft_pad(N - m)] = cplx<T>{xr + yi, -xi + yr};
        }
    }
    __syncthreads();
    if constexpr (WAVE) {
        // wavefront q transforms field f0 + q and stores the grid rows from its last stage
        const int q = tid >> 6, f = f0 + q;
        if (f >= nf) return;
        const T s = w_s, c = w_c;
        const long gf = w_g;  // grid field the transform is written to
        T *gn = grid + gf * gstride + (long)j * nlon, *gs = grid + gf * gstride + (long)j2 * nlon;
        auto none = [](int, int) { return cplx<T>{T(0), T(0)}; };
        constexpr RadPlan P = rad_plan(NC);
        if constexpr (QUADL) {
            cplx<T> *bq = buf + q * NP;
            auto lds_st = [&](int, int p, const cplx<T> &v) { bq[p] = v; };
            QuadTw<T, RL, NSL> tl;
            fft_stages_w<T, NC, 0, 1, false, false, P.n - 1>(bq, GTW ? tw : tws, +1, none, lds_st,
                                                             StageTw<T, P.r[0], NC, 1>{}, &tl);
            // the last stage: butterflies j = 4 lane + bb, outputs x = j + r NSL
            const int lane = tid & 63;
            cplx<T> v[4][RL];
#pragma unroll
            for (int bb = 0; bb < 4; bb++)
#pragma unroll
                for (int r = 0; r < RL; r++) v[bb][r] = bq[fft_pad(4 * lane + bb + r * NSL)];
#pragma unroll
            for (int bb = 0; bb < 4; bb++) {
#pragma unroll
                for (int r = 1; r < RL; r++) {
                    cplx<T> t = tl.t[bb][r - 1];
                    t.y = -t.y;  // sign +1
                    v[bb][r] = cmul(v[bb][r], t);
                }
                small_dft<T, RL>(v[bb], +1);
            }
#pragma unroll
            for (int r = 0; r < RL; r++) {
                const int x = 4 * lane + r * NSL;
                if constexpr (sizeof(T) == 4) {
                    *(float4 *)(gn + x) = float4{v[0][r].x * s + c, v[1][r].x * s + c, v[2][r].x * s + c,
                                                 v[3][r].x * s + c};
                    *(float4 *)(gs + x) = float4{v[0][r].y * s + c, v[1][r].y * s + c, v[2][r].y * s + c,
                                                 v[3][r].y * s + c};
                } else {
                    *(double2 *)(gn + x) = double2{v[0][r].x * s + c, v[1][r].x * s + c};
                    *(double2 *)(gn + x + 2) = double2{v[2][r].x * s + c, v[3][r].x * s + c};
                    *(double2 *)(gs + x) = double2{v[0][r].y * s + c, v[1][r].y * s + c};
                    *(double2 *)(gs + x + 2) = double2{v[2][r].y * s + c, v[3][r].y * s + c};
                }
            }
            return;
        }
        auto out = [&](int x, int, const cplx<T> &z) {
            gn[x] = z.x * s + c;
            gs[x] = z.y * s + c;
        };
        fft_stages_w<T, NC, 0, 1, false, true, P.n>(buf + q * NP, GTW ? tw : tws, +1, none, out,
                                                    StageTw<T, P.r[0], NC, 1>{});
        return;
    }
    if constexpr (NC > 0)
        fft_stages<T, NFFT, NC, 0, 1>(buf, tws, +1);
    else
        fft_run<T, NFFT, MAXN>(buf, N, rd.r, rd.n, tws, +1);
    const T rc = rcos[j];
#pragma unroll
    for (int q = 0; q < NFFT; q++) {
        const int f = f0 + q;
        if (f >= nf) break;
        bool dxf;
        const int lf = lfield(q, dxf);  // the mean belongs to the Legendre field; an x-gradient has none
        const T s = rflag[f] ? rc : T(1);
        const T c = (mean && !dxf) ? mean[lf] : T(0);
        const long gf = slot ? slot[f] : f;  // grid field the transform is written to
        T *gn = grid + gf * gstride + (long)j * nlon, *gs = grid + gf * gstride + (long)j2 * nlon;
        const cplx<T> *bq = buf + q * NP;
        for (int i = tid; i < N; i += 256) {
            const cplx<T> z = bq[fft_pad(i)];
            gn[i] = z.x * s + c;
            gs[i] = z.y * s + c;
        }
    }
}

// Fourier analysis: rows of a latitude pair -> weighted symmetric / antisymmetric
// Fourier coefficients, m = 0..T. Grid rows are loaded in one batch per thread.
template <typename T, int NFFT, int MAXN, int NC>
__global__ __launch_bounds__(fft_threads(NFFT, NC)) void k_fft_anal(const T *__restrict__ grid, long gstride, int nf,
                                                  const unsigned char *__restrict__ rflag,
                                                  const T *__restrict__ rcos, const T *__restrict__ wfac,
                                                  const cplx<T> *__restrict__ tw, T *__restrict__ fa, int ldf,
                                                  int nlon, int nlat, int nhp, int ntrunc, Radices rd,
                                                  const int *__restrict__ nfdyn) {
    constexpr int LN = NC ? NC : MAXN;
    constexpr bool WAVE = fft_wave(NFFT, NC), GTW = WAVE && NFFT == 8;
    constexpr int NT = fft_threads(NFFT, NC);
    __shared__ cplx<T> buf[NFFT * fft_pitch(LN)];
    __shared__ cplx<T> tws[GTW ? 1 : LN];
    const int nfg = (nf + NFFT - 1) / NFFT, tile = xcd_tile(blockIdx.x, gridDim.x);
    const int jp = tile / nfg, f0 = (tile - jp * nfg) * NFFT;
    if (nfdyn) {
        nf = min(nf, *nfdyn);
        if (f0 >= nf) return;  // a field group past the device field count
    }
    const int j = jp, j2 = nlat - 1 - jp, N = NC ? NC : nlon, NP = fft_pitch(N), tid = threadIdx.x;
    const T wf_j = wfac[j], rc_j = rcos[j];  // the output weights, loaded with the grid rows
    if constexpr (WAVE) {
        // wavefront q transforms field f0 + q, its first stage loading the grid rows;
        // a field past nf repeats field nf - 1 and is not stored
        constexpr RadPlan P = rad_plan(NC);
        const int q = tid >> 6, f = min(f0 + q, nf - 1);
        const T *gn = grid + f * gstride + (long)j * nlon, *gs = grid + f * gstride + (long)j2 * nlon;
        cplx<T> *bq = buf + q * NP;
        if constexpr (!GTW)
            for (int t = tid; t < N; t += NT) tws[t] = tw[t];
        auto in = [&](int x, int) { return cplx<T>{gn[x], gs[x]}; };
        auto lds_st = [&](int, int p, const cplx<T> &v) { bq[p] = v; };
        // the second stage's twiddles load during the first stage from the global
        // table; from the LDS table only after the barrier that completes it
        StageTw<T, P.r[1], NC, P.r[0]> t1;
        NoTw none;
        if constexpr (GTW)
            fft_stage_w<T, P.r[0], NC, 1>(tw, -1, in, lds_st, StageTw<T, P.r[0], NC, 1>{}, t1);
        else
            fft_stage_w<T, P.r[0], NC, 1>(tws, -1, in, lds_st, StageTw<T, P.r[0], NC, 1>{}, none);
        if constexpr (!GTW) {
            __syncthreads();  // the twiddle table (the global one needs no barrier)
            t1.load(tws);
        }
        fft_stages_w<T, NC, 1, P.r[0], false, false, P.n>(bq, GTW ? tw : tws, -1, in, lds_st, t1);
        __syncthreads();  // the coefficients below are read across fields
    } else {
        constexpr int MAXG = (NFFT * LN + 255) / 256;
        T xv[MAXG], yv[MAXG];
#pragma unroll
        for (int c = 0; c < MAXG; c++) {
            int t = tid + c * 256;
            xv[c] = T(0);
            yv[c] = T(0);
            if (t < NFFT * N) {
                int q = t / N, i = t - q * N, f = f0 + q;
                if (f < nf) {
                    xv[c] = grid[f * gstride + (long)j * nlon + i];
                    yv[c] = grid[f * gstride + (long)j2 * nlon + i];
                }
            }
        }
        for (int t = tid; t < N; t += 256) tws[t] = tw[t];
#pragma unroll
        for (int c = 0; c < MAXG; c++) {
            int t = tid + c * 256;
            if (t < NFFT * N) {
                const int q = t / N;
                buf[q * NP + fft_pad(t - q * N)] = cplx<T>{xv[c], yv[c]};
            }
        }
        __syncthreads();
        if constexpr (NC > 0)
            fft_stages<T, NFFT, NC, 0, 1>(buf, tws, -1);
        else
            fft_run<T, NFFT, MAXN>(buf, N, rd.r, rd.n, tws, -1);
    }
    // a thread's field is the same in every iteration (NT is a multiple of NFFT):
    // its weight is formed once
    static_assert(NT % NFFT == 0, "output field per thread");
    const int qo = tid % NFFT, fo = f0 + qo;
    const T wf = rflag[min(fo, nf - 1)] ? wf_j * rc_j : wf_j;
    for (int idx = tid; idx < (ntrunc + 1) * NFFT; idx += NT) {
        int m = idx / NFFT, q = qo, f = fo;
        if (f >= nf) continue;
        cplx<T> zk = buf[q * NP + fft_pad(m)], zn = buf[q * NP + fft_pad(m == 0 ? 0 : N - m)];
        T xr = T(0.5) * (zk.x + zn.x), xi = T(0.5) * (zk.y - zn.y);
        T yr = T(0.5) * (zk.y + zn.y), yi = T(-0.5) * (zk.x - zn.x);
        T *pe = fa + ((long)(2 * m) * nhp + jp) * ldf + 2 * f;
        T *po = fa + ((long)(2 * m + 1) * nhp + jp) * ldf + 2 * f;
        *(cplx<T> *)pe = cplx<T>{wf * (xr + yr), wf * (xi + yi)};
        *(cplx<T> *)po = cplx<T>{wf * (xr - yr), wf * (xi - yi)};
    }
}

// ---------------------------------------------------------------------------
// Gather into extended spectral columns and scatter out of them.
constexpr int GATHER_KE = 8;
constexpr int GATHER_QB = 4;  // rows whose loads are issued together

// Each output element is a combination of at most three source coefficients
// ("taps") of the row's order m: U-hat takes S at n - 1 and n + 1 and T at n,
// V-hat takes S at n and T at n - 1 and n + 1, a copy takes its field at n.
// The taps of GATHER_QB rows are loaded before any is used, with indices
// clamped into the state and the values of taps outside m <= n <= ntrunc (or
// of a missing field) zeroed afterwards, so that no load waits for another.
template <typename T>
__global__ void k_synth_gather(const cplx<T> *__restrict__ src, long sf, long sk, const SynthDesc *__restrict__ desc,
                               int nfo, T *__restrict__ sx, int ldx, const int2 *__restrict__ kemn,
                               const int *__restrict__ koff, const T *__restrict__ eps, const int *__restrict__ eoff,
                               int ntrunc, int nke, T *__restrict__ mean, const int *__restrict__ nfdyn) {
    const int fo = blockIdx.x * blockDim.x + threadIdx.x;
    if (nfdyn) nfo = min(nfo, *nfdyn);
    if (fo >= nfo) return;
    const SynthDesc d = desc[fo];
    if (mean && blockIdx.y == 0) {
        // global mean of an SK_COPY_MEAN field (Y_0^0 = 1 / sqrt(4 pi)); zero for the others
        T c = T(0);
        if (d.kind == SK_COPY_MEAN) c = T((double)src[(long)d.a * sf].x * d.sa * 0.28209479177387814);
        mean[fo] = c;
    }
    const bool uh = d.kind == SK_UHAT, vh = d.kind == SK_VHAT, vec = uh || vh;
    const bool dev = d.kind == SK_COPY_DEV || d.kind == SK_COPY_MEAN;  // without the n = 0 coefficient
    const bool none = !(vec || dev || d.kind == SK_COPY);
    // taps: field, degree offset, multiplier, invlap flag
    const int fa = max(d.a, 0), fb = max(d.b, 0);  // a missing field (-1) reads field 0, zeroed below
    const int tf0 = fa, td0 = uh ? -1 : 0;
    const int tf1 = uh ? fa : fb, td1 = uh ? 1 : -1;
    const int tf2 = fb, td2 = uh ? 0 : 1;
    const bool okA = d.a >= 0, okB = d.b >= 0;
    const bool ok0 = okA && !none, ok1 = uh ? okA : okB, ok2 = okB;
    const T sa = T(d.sa), sb = T(d.sb);
    const T s0 = sa, s1 = uh ? sa : sb, s2 = sb;
    const int i0 = d.inv_a, i1 = uh ? d.inv_a : d.inv_b, i2 = d.inv_b;
    auto tap = [&](int f, int m, int n) {
        const int nc = min(max(n, m), ntrunc);
        return src[f * sf + (long)(koff[m] + nc - m) * sk];
    };
    auto fac = [&](bool ok, int m, int n, T sc, int inv) {
        if (!ok || n < m || n > ntrunc) return T(0);
        if (inv) sc *= (n == 0) ? T(0) : T(-1.0 / ((double)n * (n + 1.0)));
        return sc;
    };
    const int kb0 = blockIdx.y * GATHER_KE, kb1 = min(nke, kb0 + GATHER_KE);
    for (int kb = kb0; kb < kb1; kb += GATHER_QB) {
        int2 mn[GATHER_QB];
        cplx<T> v0[GATHER_QB], v1[GATHER_QB], v2[GATHER_QB];
#pragma unroll
        for (int q = 0; q < GATHER_QB; q++) {
            mn[q] = kemn[min(kb + q, kb1 - 1)];  // clamped: rows past kb1 are not stored
            const int m = mn[q].x, n = mn[q].y;
            v0[q] = tap(tf0, m, n + td0);
            if (vec) {
                v1[q] = tap(tf1, m, n + td1);
                v2[q] = tap(tf2, m, n + td2);
            }
        }
#pragma unroll
        for (int q = 0; q < GATHER_QB; q++) {
            const int ke = kb + q;
            if (ke >= kb1) break;
            const int m = mn[q].x, n = mn[q].y;
            cplx<T> r{T(0), T(0)};
            if (vec) {
                const T *ep = eps + eoff[m];
                const T en = ep[n - m], en1 = ep[n + 1 - m];
                const T cm1 = T(n - 1) * en, cp1 = T(n + 2) * en1, fm = T(m);
                const T c0 = fac(ok0, m, n + td0, s0, i0), c1 = fac(ok1, m, n + td1, s1, i1);
                const T c2 = fac(ok2, m, n + td2, s2, i2);
                const cplx<T> a0{v0[q].x * c0, v0[q].y * c0}, a1{v1[q].x * c1, v1[q].y * c1};
                const cplx<T> a2{v2[q].x * c2, v2[q].y * c2};
                if (uh) {
                    // U = (n-1) eps_n S_{n-1} - (n+2) eps_{n+1} S_{n+1} + i m T_n
                    r.x = cm1 * a0.x - cp1 * a1.x - fm * a2.y;
                    r.y = cm1 * a0.y - cp1 * a1.y + fm * a2.x;
                } else {
                    // V = i m S_n - (n-1) eps_n T_{n-1} + (n+2) eps_{n+1} T_{n+1}
                    r.x = -fm * a0.y - cm1 * a1.x + cp1 * a2.x;
                    r.y = fm * a0.x - cm1 * a1.y + cp1 * a2.y;
                }
            } else if (!none && (!dev || n > 0)) {
                const T c0 = fac(ok0, m, n, s0, i0);
                r = cplx<T>{v0[q].x * c0, v0[q].y * c0};
            }
            *(cplx<T> *)(sx + (long)ke * ldx + 2 * fo) = r;
        }
    }
}

// SCATTER_KE coefficients per workgroup (one descriptor load, SCATTER_KE rows),
// threads over the output fields.
constexpr int SCATTER_KE = 8;

// MFX: the descriptors may hold AK_MASSFIX (a separate instantiation keeps its
// division out of the plain scatter)
template <typename T, bool MFX>
__global__ void k_anal_scatter(const T *__restrict__ sx, int ldx, const AnalDesc *__restrict__ desc, int nfo,
                               cplx<T> *__restrict__ dst, long df, long dk, const int2 *__restrict__ kmn,
                               const int *__restrict__ keoff, const T *__restrict__ eps,
                               const int *__restrict__ eoff, T rrs, int nlm, const int *__restrict__ nfdyn) {
    const int fo = blockIdx.x * blockDim.x + threadIdx.x;
    if (fo >= nfo) return;
    const AnalDesc d = desc[fo];
    const int k0 = blockIdx.y * SCATTER_KE, kend = min(nlm, k0 + SCATTER_KE);
    // a source field past the device field count was not transformed (the
    // caller has established that it is zero on the grid): its output is zero
    const int lim = nfdyn ? *nfdyn : INT_MAX;
    if (d.kind == AK_ZERO || d.a >= lim || d.b >= lim || d.c >= lim) {
        if (!d.accum)
            for (int k = k0; k < kend; k++) dst[fo * df + (long)k * dk] = cplx<T>{T(0), T(0)};
        return;
    }
    // the inputs of QB rows first (loads all in flight together), then those rows;
    // FP64 takes half the rows per batch to stay within the registers
    constexpr int QB = sizeof(T) == 4 ? SCATTER_KE : SCATTER_KE / 2;
    const bool mfx = MFX && d.kind == AK_MASSFIX, copy = d.kind == AK_COPY || mfx, rd = d.accum || mfx;
    const int ca = (d.kind == AK_VRT) ? d.b : d.a;  // i m (V for vrt, U for div; the field for a copy)
    const int cb = (d.kind == AK_VRT) ? d.a : d.b;  // recurrence partner (U for vrt, V for div)
    for (int kb = k0; kb < kend; kb += QB) {
        cplx<T> z0[QB], zp[QB], zm[QB], zc[QB], za[QB];
        int2 mnq[QB];
#pragma unroll
        for (int q = 0; q < QB; q++) {
            const int k = min(kb + q, kend - 1);  // clamped: rows past kend are not stored
            const int2 mn = kmn[k];
            const int m = mn.x, n = mn.y;
            mnq[q] = mn;
            const T *X = sx + (long)(keoff[m] + n - m) * ldx;
            const T *C = copy ? X + 2 * d.a : X + 2 * ca;
            z0[q] = cplx<T>{C[0], C[1]};
            zp[q] = zm[q] = zc[q] = za[q] = cplx<T>{T(0), T(0)};
            if (!copy) {
                zp[q] = cplx<T>{X[ldx + 2 * cb], X[ldx + 2 * cb + 1]};
                if (n > m) zm[q] = cplx<T>{X[-ldx + 2 * cb], X[-ldx + 2 * cb + 1]};
            }
            if (d.c >= 0) zc[q] = cplx<T>{X[2 * d.c], X[2 * d.c + 1]};
            if (rd) za[q] = dst[fo * df + (long)k * dk];
        }
#pragma unroll
        for (int q = 0; q < QB; q++) {
            const int k = kb + q;
            if (k >= kend) break;
            const int m = mnq[q].x, n = mnq[q].y;
            cplx<T> r = z0[q];
            if (!copy) {
                const T *ep = eps + eoff[m];
                T en = ep[n - m], en1 = ep[n + 1 - m];
                T cp = T(n) * en1, cm = T(n + 1) * en, fm = T(m);
                T sgn = (d.kind == AK_VRT) ? T(-1) : T(1);
                // vrt = (i m V_n - n eps_{n+1} U_{n+1} + (n+1) eps_n U_{n-1}) / a
                // div = (i m U_n + n eps_{n+1} V_{n+1} - (n+1) eps_n V_{n-1}) / a
                r.x = (-fm * z0[q].y + sgn * (cp * zp[q].x - cm * zm[q].x)) * rrs;
                r.y = (fm * z0[q].x + sgn * (cp * zp[q].y - cm * zm[q].y)) * rrs;
            }
            if (d.c >= 0) {
                T qq = T(d.sc) * T(n) * T(n + 1);
                r.x += qq * zc[q].x;
                r.y += qq * zc[q].y;
            }
            T sc = T(d.scale);
            cplx<T> *o = dst + fo * df + (long)k * dk;
            if (MFX && mfx) {
                // ln ps + dt (t - ln ps) / dt, t the analysed deviation plus the mean at n = 0
                const T dt = sc;
                const cplx<T> s = za[q];
                cplx<T> t = r;
                if (k == 0) t.x += s.x;
                const cplx<T> dd{(t.x - s.x) / dt, (t.y - s.y) / dt};
                *o = cplx<T>{s.x + dt * dd.x, s.y + dt * dd.y};
                continue;
            }
            if (d.inv) sc *= (n == 0) ? T(0) : T(-1.0 / ((double)n * (n + 1.0)));
            if (d.accum)
                *o = cplx<T>{za[q].x + sc * r.x, za[q].y + sc * r.y};
            else
                *o = cplx<T>{sc * r.x, sc * r.y};
        }
    }
}

// ---------------------------------------------------------------------------
// Plan construction

static Radices factor_nlon(int n) {
    Radices r{};
    r.n = 0;
    while (n % 4 == 0 && n > 4) {
        r.r[r.n++] = 4;
        n /= 4;
    }
    while (n % 3 == 0) {
        r.r[r.n++] = 3;
        n /= 3;
    }
    while (n % 4 == 0) {
        r.r[r.n++] = 4;
        n /= 4;
    }
    while (n % 2 == 0) {
        r.r[r.n++] = 2;
        n /= 2;
    }
    if (n != 1) r.n = -1;
    return r;
}

template <typename T> static int upload(T **d, const void *h, size_t bytes, hipStream_t s) {
    GFSDC_CHECK(hipMalloc((void **)d, bytes > 0 ? bytes : 1));
    if (bytes) GFSDC_CHECK(hipMemcpyAsync(*d, h, bytes, hipMemcpyHostToDevice, s));
    return 0;
}

template <typename T> int plan_create(Plan<T> &p, int nlon, int nlat, int ntrunc, hipStream_t s) {
    if (nlon <= 0 || nlat <= 1 || ntrunc < 1 || nlat % 2 || 2 * ntrunc >= nlon || nlat < ntrunc + 1) {
        set_error("plan: need even nlat >= ntrunc+1 and nlon > 2*ntrunc");
        return GFSDC_EINVAL;
    }
    Radices rd = factor_nlon(nlon);
    if (rd.n <= 0 || nlon > FftCfg<T>::MAXN) {
        set_error("plan: nlon must be 2^a 3^b and <= 768");
        return GFSDC_EINVAL;
    }
    p.nlon = nlon;
    p.nlat = nlat;
    p.nhalf = nlat / 2;
    p.nhp = round_up(p.nhalf, GEMM_BM);
    p.ntrunc = ntrunc;
    p.nlm = (ntrunc + 1) * (ntrunc + 2) / 2;
    p.nlme = p.nlm + ntrunc + 1;
    p.nrad = rd.n;
    for (int i = 0; i < rd.n; i++) p.rad[i] = rd.r[i];
    p.fft_nf = FftCfg<T>::NFFT;
    if (const char *e = getenv("GFSDC_FFT_NF")) {
        int v = atoi(e);
        if (v == 2 || v == 4 || v == 8) p.fft_nf = v;
    }
    p.fft_nf_an = p.fft_nf;
    const int T_ = ntrunc;
    // offsets
    p.poff_h.resize(T_ + 2);
    p.keoff_h.resize(T_ + 2);
    p.koff_h.resize(T_ + 2);
    std::vector<int> eoff_h(T_ + 2);
    long po = 0;
    int ke = 0, k = 0, eo = 0;
    for (int m = 0; m <= T_ + 1; m++) {
        p.poff_h[m] = po;
        p.keoff_h[m] = ke;
        p.koff_h[m] = k;
        eoff_h[m] = eo;
        if (m <= T_) {
            po += (long)(T_ + 2 - m) * p.nhp;
            ke += T_ + 2 - m;
            k += T_ + 1 - m;
            eo += T_ + 3 - m;
        }
    }
    std::vector<T> epsv(eo);
    for (int m = 0; m <= T_; m++)
        for (int n = m; n <= T_ + 2; n++) {
            double e = (n <= m) ? 0.0 : std::sqrt(((double)n * n - (double)m * m) / (4.0 * n * n - 1.0));
            epsv[eoff_h[m] + n - m] = (T)e;
        }
    std::vector<int2> kemn(p.nlme), kmn(p.nlm);
    ke = 0;
    k = 0;
    for (int m = 0; m <= T_; m++)
        for (int n = m; n <= T_ + 1; n++) {
            kemn[ke++] = int2{m, n};
            if (n <= T_) kmn[k++] = int2{m, n};
        }
    // analysis m-tiles: rows [i0, i0+BM) of the even block of order m
    std::vector<int> mt_m, mt_i;
    for (int m = 0; m <= T_; m++) {
        int Ke = (T_ + 3 - m) / 2;
        for (int i0 = 0; i0 < Ke; i0 += GEMM_BM) {
            mt_m.push_back(m);
            mt_i.push_back(i0);
        }
    }
    p.n_mtiles = (int)mt_m.size();
    int rc;
    if ((rc = upload(&p.poff, p.poff_h.data(), sizeof(long) * (T_ + 2), s))) return rc;
    if ((rc = upload(&p.keoff, p.keoff_h.data(), sizeof(int) * (T_ + 2), s))) return rc;
    if ((rc = upload(&p.koff, p.koff_h.data(), sizeof(int) * (T_ + 2), s))) return rc;
    if ((rc = upload(&p.eoff, eoff_h.data(), sizeof(int) * (T_ + 2), s))) return rc;
    if ((rc = upload(&p.eps, epsv.data(), sizeof(T) * eo, s))) return rc;
    if ((rc = upload(&p.mtile_m, mt_m.data(), sizeof(int) * mt_m.size(), s))) return rc;
    if ((rc = upload(&p.mtile_i, mt_i.data(), sizeof(int) * mt_i.size(), s))) return rc;
    if ((rc = upload(&p.kemn, kemn.data(), sizeof(int2) * kemn.size(), s))) return rc;
    if ((rc = upload(&p.kmn, kmn.data(), sizeof(int2) * kmn.size(), s))) return rc;
    // Gauss nodes and Legendre tables in FP64 on the device
    GFSDC_CHECK(hipMalloc((void **)&p.mu_d, sizeof(double) * nlat));
    GFSDC_CHECK(hipMalloc((void **)&p.w_d, sizeof(double) * nlat));
    hipLaunchKernelGGL(k_gauss, dim3((nlat / 2 + 63) / 64), dim3(64), 0, s, p.mu_d, p.w_d, nlat);
    GFSDC_CHECK(hipGetLastError());
    GFSDC_CHECK(hipMalloc((void **)&p.plm, sizeof(T) * (size_t)po));
    hipLaunchKernelGGL(k_legendre_table<T>, dim3(p.nhp / 64, T_ + 1), dim3(64), 0, s, p.plm, p.poff, p.mu_d,
                       p.nhalf, p.nhp, T_);
    GFSDC_CHECK(hipGetLastError());
    // latitude-major copy for the analysis: per order, [parity][j][i'], rows of
    // Kw = roundup(Ke, GEMM_BM) degrees, zero past each parity's last degree
    std::vector<long> tpoff_h(T_ + 2);
    long to = 0;
    for (int m = 0; m <= T_ + 1; m++) {
        tpoff_h[m] = to;
        if (m <= T_) to += 2L * p.nhp * round_up((T_ + 3 - m) / 2, GEMM_BM);
    }
    if ((rc = upload(&p.tpoff, tpoff_h.data(), sizeof(long) * (T_ + 2), s))) return rc;
    GFSDC_CHECK(hipMalloc((void **)&p.plmt, sizeof(T) * (size_t)to));
    hipLaunchKernelGGL(k_transpose_plm<T>, dim3(p.nhp, 2, T_ + 1), dim3(GEMM_BM), 0, s, p.plm, p.poff, p.plmt, p.tpoff,
                       p.nhp, T_);
    GFSDC_CHECK(hipGetLastError());
    p.mu.resize(nlat);
    p.w.resize(nlat);
    GFSDC_CHECK(hipMemcpyAsync(p.mu.data(), p.mu_d, sizeof(double) * nlat, hipMemcpyDeviceToHost, s));
    GFSDC_CHECK(hipMemcpyAsync(p.w.data(), p.w_d, sizeof(double) * nlat, hipMemcpyDeviceToHost, s));
    GFSDC_CHECK(hipStreamSynchronize(s));
    std::vector<T> rcos(nlat), wfac(nlat), tw(2 * nlon);
    for (int j = 0; j < nlat; j++) {
        rcos[j] = (T)(1.0 / std::sqrt((1.0 - p.mu[j]) * (1.0 + p.mu[j])));
        wfac[j] = (T)(2.0 * M_PI * p.w[j] / nlon);
    }
    // twiddles in stage order: the stage with radix R after Ns = (product of the
    // earlier radices) holds w^(jm r nlon / (Ns R)), w = exp(-2 pi i / nlon), at
    // Ns - 1 + (r - 1) Ns + jm for r = 1..R-1, jm = 0..Ns-1 (nlon - 1 entries in
    // all), so the lanes of a stage, which have consecutive jm, read
    // consecutive entries instead of a stride of r nlon / (Ns R).
    std::fill(tw.begin(), tw.end(), T(0));
    for (int st = 0, Ns = 1; st < rd.n; Ns *= rd.r[st], st++) {
        const int R = rd.r[st], tstep = nlon / (Ns * R);
        for (int r = 1; r < R; r++)
            for (int jm = 0; jm < Ns; jm++) {
                const int e = Ns - 1 + (r - 1) * Ns + jm;
                const double a = 2.0 * M_PI * (double)((long)jm * r * tstep) / nlon;
                tw[2 * e] = (T)std::cos(a);
                tw[2 * e + 1] = (T)(-std::sin(a));
            }
    }
    if ((rc = upload(&p.rcos, rcos.data(), sizeof(T) * nlat, s))) return rc;
    if ((rc = upload(&p.wfac, wfac.data(), sizeof(T) * nlat, s))) return rc;
    if ((rc = upload(&p.tw, tw.data(), sizeof(T) * 2 * nlon, s))) return rc;
    GFSDC_CHECK(hipStreamSynchronize(s));
    return 0;
}

template <typename T> void plan_destroy(Plan<T> &p) {
    void *ptrs[] = {p.plm, p.poff, p.plmt, p.tpoff, p.keoff, p.koff, p.eps, p.eoff, p.rcos, p.wfac,
                    p.tw, p.mtile_m, p.mtile_i, p.mu_d, p.w_d, p.kemn, p.kmn};
    for (void *q : ptrs)
        if (q) (void)hipFree(q);
    p = Plan<T>();
}

template <typename T> int work_alloc(Work<T> &w, const Plan<T> &p, int nfield) {
    if (nfield <= w.ncap) return 0;
    work_free(w);
    w.ncap = nfield;
    w.ldc = round_up(2 * nfield, COL_PAD) + COL_SKEW;
    size_t nsx = (size_t)p.nlme * w.ldc, nfs = (size_t)(p.ntrunc + 1) * p.nlat * w.ldc;
    size_t nfa = (size_t)(p.ntrunc + 1) * 2 * p.nhp * w.ldc, ng = (size_t)nfield * p.nlat * p.nlon;
    GFSDC_CHECK(hipMalloc((void **)&w.sx, sizeof(T) * nsx));
    GFSDC_CHECK(hipMalloc((void **)&w.fs, sizeof(T) * nfs));
    GFSDC_CHECK(hipMalloc((void **)&w.fa, sizeof(T) * nfa));
    GFSDC_CHECK(hipMalloc((void **)&w.grid, sizeof(T) * ng));
    GFSDC_CHECK(hipMemset(w.sx, 0, sizeof(T) * nsx));
    GFSDC_CHECK(hipMemset(w.fs, 0, sizeof(T) * nfs));
    GFSDC_CHECK(hipMemset(w.fa, 0, sizeof(T) * nfa));
    return 0;
}

template <typename T> void work_free(Work<T> &w) {
    if (w.sx) (void)hipFree(w.sx);
    if (w.fs) (void)hipFree(w.fs);
    if (w.fa) (void)hipFree(w.fa);
    if (w.grid) (void)hipFree(w.grid);
    w = Work<T>();
}

// ---------------------------------------------------------------------------
// launchers

template <typename T>
void launch_synth_gather(const Plan<T> &p, const cplx<T> *src, long sf, long sk, const SynthDesc *desc, int nfo,
                         T *sx, int ldx, double rsphere, hipStream_t s, T *mean, const int *nfdyn) {
    (void)rsphere;
    dim3 grid((nfo + 255) / 256, (p.nlme + GATHER_KE - 1) / GATHER_KE);
    hipLaunchKernelGGL(k_synth_gather<T>, grid, dim3(256), 0, s, src, sf, sk, desc, nfo, sx, ldx, p.kemn, p.koff,
                       p.eps, p.eoff, p.ntrunc, p.nlme, mean, nfdyn);
}

// The BF16x3 Legendre kernels take no device field count: they transform every
// column of the launch (the columns past it hold earlier data and are not read).
template <typename T>
void launch_leg_synth(const Plan<T> &p, const T *sx, int ldx, T *fs, int ldf, int ncol, hipStream_t s,
                      const int *nfdyn) {
    int ntj = p.nhp / GEMM_BM, ntc = (ncol + GT<T>::BN - 1) / GT<T>::BN;
    if constexpr (sizeof(T) == 4) {
        if (p.leg_bf16x3) {
            hipLaunchKernelGGL(k_leg_synth_b3, dim3((ncol + GEMM_BN - 1) / GEMM_BN, (p.ntrunc + 1) * ntj), dim3(256), 0,
                               s, p.plm, p.poff, p.keoff, sx, ldx, fs, ldf, p.nhalf, p.nhp, p.nlat, p.ntrunc, ntj);
            return;
        }
    }
    hipLaunchKernelGGL(k_leg_synth<T>, dim3(ntc * 2 * (p.ntrunc + 1) * ntj), dim3(256), 0, s, p.plm, p.poff,
                       p.keoff, sx, ldx, fs, ldf, p.nhalf, p.nhp, p.nlat, p.ntrunc, ntj, ncol, nfdyn);
}

template <typename T>
void launch_leg_anal(const Plan<T> &p, const T *fa, int ldf, T *sx, int ldx, int ncol, hipStream_t s,
                     const int *nfdyn) {
    int ntc = (ncol + GT<T>::BN - 1) / GT<T>::BN;
    if constexpr (sizeof(T) == 4) {
        if (p.leg_bf16x3) {
            hipLaunchKernelGGL(k_leg_anal_b3, dim3((ncol + GEMM_BN - 1) / GEMM_BN, p.n_mtiles), dim3(256), 0, s, p.plm,
                               p.poff, p.keoff, p.mtile_m, p.mtile_i, fa, ldf, sx, ldx, p.nhp, p.ntrunc);
            return;
        }
    }
    hipLaunchKernelGGL(k_leg_anal<T>, dim3(ntc * p.n_mtiles * 2), dim3(256), 0, s, p.plmt, p.tpoff, p.keoff,
                       p.mtile_m, p.mtile_i, fa, ldf, sx, ldx, p.nhp, p.ntrunc, ncol, p.n_mtiles, nfdyn);
}

static Radices radices_of(int n, const int *r) {
    Radices rd{};
    rd.n = n;
    for (int i = 0; i < n; i++) rd.r[i] = r[i];
    return rd;
}

template <typename T, int NF>
static void fft_synth_nf(const Plan<T> &p, const T *fs, int ldf, T *grid, long gstride, int nf,
                         const unsigned char *rflag, const T *mean, const int *slot, int fofs, hipStream_t s,
                         const int *nfdyn, const FieldMap &fmap, T rdx) {
    constexpr int MN = FftCfg<T>::MAXN;
    dim3 g(p.nhalf * ((nf + NF - 1) / NF));
#define FFT_SYN(NC)                                                                                             \
    hipLaunchKernelGGL((k_fft_synth<T, NF, MN, NC>), g, dim3(fft_threads(NF, NC)), 0, s, fs, ldf, grid, gstride, nf,   \
                       rflag, mean, slot,                                                                     \
                       p.rcos, (const cplx<T> *)p.tw, p.nlon, p.nlat, p.ntrunc, radices_of(p.nrad, p.rad), fofs, \
                       nfdyn, fmap, rdx)
    switch (p.nlon) {
    case 768: FFT_SYN(768); break;
    case 384: FFT_SYN(384); break;
    case 128: FFT_SYN(128); break;
    default: FFT_SYN(0); break;
    }
#undef FFT_SYN
}

template <typename T, int NF>
static void fft_anal_nf(const Plan<T> &p, const T *grid, long gstride, int nf, const unsigned char *rflag, T *fa,
                        int ldf, hipStream_t s, const int *nfdyn) {
    constexpr int MN = FftCfg<T>::MAXN;
    dim3 g(p.nhalf * ((nf + NF - 1) / NF));
#define FFT_ANA(NC)                                                                                             \
    hipLaunchKernelGGL((k_fft_anal<T, NF, MN, NC>), g, dim3(fft_threads(NF, NC)), 0, s, grid, gstride, nf, rflag,    \
                       p.rcos, p.wfac,                                                                        \
                       (const cplx<T> *)p.tw, fa, ldf, p.nlon, p.nlat, p.nhp, p.ntrunc, radices_of(p.nrad, p.rad), \
                       nfdyn)
    switch (p.nlon) {
    case 768: FFT_ANA(768); break;
    case 384: FFT_ANA(384); break;
    case 128: FFT_ANA(128); break;
    default: FFT_ANA(0); break;
    }
#undef FFT_ANA
}

// fields per workgroup of the FFT kernels (p.fft_nf for the synthesis and
// p.fft_nf_an for the analysis, chosen by the tile autotuner or GFSDC_FFT_NF):
// more fields give longer contiguous Fourier segments, fewer give more
// workgroups per CU (LDS = fields * nlon * sizeof(complex))
template <typename T>
void launch_fft_synth(const Plan<T> &p, const T *fs, int ldf, T *grid, long gstride, int nf,
                      const unsigned char *rflag, hipStream_t s, const T *mean, const int *slot, int fofs,
                      const int *nfdyn, const FieldMap *fmap, double rdx) {
    const FieldMap id{}, &fm = fmap ? *fmap : id;
    switch (p.fft_nf) {
    case 2: fft_synth_nf<T, 2>(p, fs, ldf, grid, gstride, nf, rflag, mean, slot, fofs, s, nfdyn, fm, (T)rdx); break;
    case 4: fft_synth_nf<T, 4>(p, fs, ldf, grid, gstride, nf, rflag, mean, slot, fofs, s, nfdyn, fm, (T)rdx); break;
    default: fft_synth_nf<T, 8>(p, fs, ldf, grid, gstride, nf, rflag, mean, slot, fofs, s, nfdyn, fm, (T)rdx); break;
    }
}

template <typename T>
void launch_fft_anal(const Plan<T> &p, const T *grid, long gstride, int nf, const unsigned char *rflag, T *fa,
                     int ldf, hipStream_t s, const int *nfdyn) {
    switch (p.fft_nf_an) {
    case 2: fft_anal_nf<T, 2>(p, grid, gstride, nf, rflag, fa, ldf, s, nfdyn); break;
    case 4: fft_anal_nf<T, 4>(p, grid, gstride, nf, rflag, fa, ldf, s, nfdyn); break;
    default: fft_anal_nf<T, 8>(p, grid, gstride, nf, rflag, fa, ldf, s, nfdyn); break;
    }
}

template <typename T>
void launch_anal_scatter(const Plan<T> &p, const T *sx, int ldx, const AnalDesc *desc, int nfo, cplx<T> *dst,
                         long df, long dk, double rsphere, hipStream_t s, const int *nfdyn, bool massfix) {
    const int bs = nfo >= 256 ? 256 : 64 * ((nfo + 63) / 64);
    dim3 grid((nfo + bs - 1) / bs, (p.nlm + SCATTER_KE - 1) / SCATTER_KE);
    auto kern = massfix ? k_anal_scatter<T, true> : k_anal_scatter<T, false>;
    hipLaunchKernelGGL(kern, grid, dim3(bs), 0, s, sx, ldx, desc, nfo, dst, df, dk, p.kmn, p.keoff, p.eps, p.eoff,
                       (T)(1.0 / rsphere), p.nlm, nfdyn);
}

#define INST(T)                                                                                                \
    template int plan_create<T>(Plan<T> &, int, int, int, hipStream_t);                                        \
    template void plan_destroy<T>(Plan<T> &);                                                                  \
    template int work_alloc<T>(Work<T> &, const Plan<T> &, int);                                               \
    template void work_free<T>(Work<T> &);                                                                     \
    template void launch_synth_gather<T>(const Plan<T> &, const cplx<T> *, long, long, const SynthDesc *, int,  \
                                         T *, int, double, hipStream_t, T *, const int *);                     \
    template void launch_leg_synth<T>(const Plan<T> &, const T *, int, T *, int, int, hipStream_t,             \
                                      const int *);                                                            \
    template void launch_leg_anal<T>(const Plan<T> &, const T *, int, T *, int, int, hipStream_t, const int *); \
    template void launch_fft_synth<T>(const Plan<T> &, const T *, int, T *, long, int, const unsigned char *,  \
                                      hipStream_t, const T *, const int *, int, const int *, const FieldMap *, \
                                      double);                                                                 \
    template void launch_fft_anal<T>(const Plan<T> &, const T *, long, int, const unsigned char *, T *, int,   \
                                     hipStream_t, const int *);                                                \
    template void launch_anal_scatter<T>(const Plan<T> &, const T *, int, const AnalDesc *, int, cplx<T> *,    \
                                         long, long, double, hipStream_t, const int *, bool);
INST(float)
INST(double)

}  // namespace gfsdc
