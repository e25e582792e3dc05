// Internal declarations of the gfsdc HIP library (gfx950 only).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/gfsdc.h"

#define GFSDC_CHECK(expr)                                                              \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess) {                                                        \
            gfsdc::set_error(std::string(#expr) + ": " + hipGetErrorString(e_));       \
            return GFSDC_EHIP;                                                         \
        }                                                                              \
    } while (0)

namespace gfsdc {

void set_error(const std::string &msg);

template <typename T> struct cplx { T x, y; };

// Workgroups are dispatched to the 8 XCDs round-robin (workgroup b on XCD b % 8).
// Give each XCD a contiguous range of logical tiles, so that tiles sharing
// cache lines (the field groups of one latitude pair; the Legendre block and
// the spectral columns of one order m; the level operators of one degree) run
// on the same XCD and meet in its L2 instead of fetching the same lines
// through eight L2s.
__device__ inline int xcd_tile(int b, int nb) {
    const int q = nb >> 3, r = nb & 7, x = b & 7, k = b >> 3;
    return x * q + min(x, r) + k;
}

// Legendre GEMM tile shape (rows x columns of one workgroup, 4 waves).
constexpr int GEMM_BM = 64;
constexpr int GEMM_BN = 128;
constexpr int GEMM_BK = 16;
// bytes of the per-device scratch of gfsdc_spectogrd_noalloc
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)128 << 20;
constexpr int COL_PAD = GEMM_BN;  // column padding of transform work buffers
// columns past the padded width: keeps the work buffers' row pitch off a power
// of two (at 1024 columns the rows that one Fourier workgroup reads or writes,
// 1.5 MiB apart, crowd onto few memory channels)
constexpr int COL_SKEW = 32;

// Synthesis Fourier buffer: for each row (E sums at row j, O sums at row
// nlat-1-j) and each group of 16 fields (32 values, one 128-byte line), the
// lines of all orders m = 0..T are contiguous, so that a Fourier workgroup's
// rows are one contiguous run per group instead of 255 scattered pieces, while
// each line is still written whole by one wavefront of the Legendre synthesis.
// Value c (2 field + re/im) of order m in row r sits at fs_index(r, m, c, ldf,
// T + 1); ldf (the work buffers' column pitch) is a multiple of 32, and the
// buffer holds nlat ldf (T+1) values.
__host__ __device__ inline long fs_index(int r, int m, int c, int ldf, int nm1) {
    return (((long)r * (ldf >> 5) + (c >> 5)) * nm1 + m) * 32 + (c & 31);
}
static_assert(COL_PAD % 32 == 0 && COL_SKEW % 32 == 0, "the Fourier buffer groups 32 columns");

inline int round_up(int a, int b) { return (a + b - 1) / b * b; }

// Transform plan: Gauss grid, Legendre tables and FFT twiddles for one
// (nlon, nlat, ntrunc) at precision T. Tables are generated on the device in FP64.
template <typename T> struct Plan {
    int nlon = 0, nlat = 0, nhalf = 0, nhp = 0, ntrunc = 0, nlm = 0, nlme = 0;
    int nrad = 0, rad[16] = {0};
    int fft_nf = 8;          // fields per synthesis FFT workgroup (2, 4 or 8)
    int fft_nf_an = 8;       // fields per analysis FFT workgroup (2, 4 or 8)
    int leg_bf16x3 = 0;      // FP32 plan only: Legendre GEMMs as three bf16 products (hi.hi + hi.lo + lo.hi)
    T *plm = nullptr;        // per m: [even i'][nhp] then [odd i'][nhp], n = m + 2i' (+1)
    T *plmt = nullptr;       // per m: [parity][nhp][Kw] latitude-major copy (analysis), Kw = roundup(Ke, GEMM_BM)
    long *tpoff = nullptr;   // device, offset of order m in plmt
    long *poff = nullptr;    // device, per m (ntrunc+2 entries)
    int *keoff = nullptr;    // device, extended spectral offset of order m
    int *koff = nullptr;     // device, spectral offset of order m
    T *eps = nullptr;        // device, eps_n^m at extended index (n = m..T+1) plus n=T+2 row
    int *eoff = nullptr;     // device, offset of order m in eps (T+3-m entries)
    T *rcos = nullptr;       // device, 1/cos(lat), nlat
    T *wfac = nullptr;       // device, 2 pi w_j / nlon, nlat
    T *tw = nullptr;         // device, complex FFT twiddles in stage order (see plan_create), 2*nlon
    int *mtile_m = nullptr;  // device, analysis m-tile -> m
    int *mtile_i = nullptr;  // device, analysis m-tile -> first row
    int n_mtiles = 0;
    int2 *kemn = nullptr;    // device, extended index -> (m, n)
    int2 *kmn = nullptr;     // device, spectral index -> (m, n)
    double *mu_d = nullptr, *w_d = nullptr;  // FP64 nodes / weights (device)
    std::vector<double> mu, w;               // host copies
    std::vector<int> keoff_h, koff_h;
    std::vector<long> poff_h;
};

template <typename T> int plan_create(Plan<T> &p, int nlon, int nlat, int ntrunc, hipStream_t s);
template <typename T> void plan_destroy(Plan<T> &p);

// Descriptors of the synthesis gather (spectral state -> extended spectral
// columns) and of the analysis scatter (extended columns -> spectral output).
// SK_COPY_DEV copies a field without its global mean (n = 0) coefficient.
// SK_COPY_MEAN does the same and hands the mean, q_00 / sqrt(4 pi), to the
// Fourier synthesis, which adds it to every grid point after the transform: the
// Legendre sums and FFT stages then round relative to the field's deviation
// (a few K for temperature) instead of its mean (300 K).
enum SynthKind : int { SK_ZERO = 0, SK_COPY = 1, SK_UHAT = 2, SK_VHAT = 3, SK_COPY_DEV = 4, SK_COPY_MEAN = 5 };
struct SynthDesc {
    int kind;
    int a, b;          // source fields (S and T of the vector pair; a for COPY)
    double sa, sb;     // multipliers of S and T
    int inv_a, inv_b;  // multiply S / T by invlap(n)
};
// Map from the grid fields of a Fourier synthesis to the Legendre-synthesised
// fields it reads: per member, per_dst grid fields in nseg runs; run s starts
// at grid field start[s] (ascending; an empty run shares its start with the
// next) and reads Legendre fields from src[s] on. A run whose bit is set in dx
// is the x-gradient (1/(a cos phi)) d/dlambda of its source: the Fourier
// synthesis multiplies the source's coefficients by i m / a instead of taking
// a Legendre synthesis of its own (the 1/cos phi comes with the field's flag).
// nseg = 0 is the identity. The member index is f * rdiv >> 32 with rdiv =
// ceil(2^32 / per_dst) (exact for f per_dst < 2^32), and the runs are tested
// with compile-time indices only, so that the kernel keeps the map in scalar
// registers and loads nothing for it.
constexpr int FMAP_SEGS = 10;
struct FieldMap {
    int nseg = 0, per_dst = 0, per_src = 0;
    unsigned rdiv = 0;
    int start[FMAP_SEGS] = {0}, src[FMAP_SEGS] = {0};
    unsigned dx = 0;
};
inline void fmap_set_divisor(FieldMap &M) {
    M.rdiv = (unsigned)((((unsigned long long)1 << 32) + (unsigned)M.per_dst - 1) / (unsigned)M.per_dst);
}
__host__ __device__ inline int fmap_field(const FieldMap &M, int f, bool &dx) {
    if (M.nseg == 0) {
        dx = false;
        return f;
    }
    const int mem = (int)(((unsigned long long)(unsigned)f * M.rdiv) >> 32), fl = f - mem * M.per_dst;
    int off = M.src[0] - M.start[0];
    bool d = M.dx & 1u;
#pragma unroll
    for (int i = 1; i < FMAP_SEGS; i++)
        if (i < M.nseg && fl >= M.start[i]) {
            off = M.src[i] - M.start[i];
            d = (M.dx >> i) & 1u;
        }
    dx = d;
    return mem * M.per_src + fl + off;
}
// grid fields fa..fb (fa <= fb) read consecutive Legendre fields of one kind:
// one member, no run starting inside
__host__ __device__ inline bool fmap_linear(const FieldMap &M, int fa, int fb) {
    if (M.nseg == 0) return true;
    const int ma = (int)(((unsigned long long)(unsigned)fa * M.rdiv) >> 32);
    const int mb = (int)(((unsigned long long)(unsigned)fb * M.rdiv) >> 32);
    if (ma != mb) return false;
    const int la = fa - ma * M.per_dst, lb = fb - ma * M.per_dst;
    bool ok = true;
#pragma unroll
    for (int i = 1; i < FMAP_SEGS; i++)
        if (i < M.nseg && M.start[i] > la && M.start[i] <= lb) ok = false;
    return ok;
}

// AK_ZERO writes a zero coefficient (an output with no source, e.g. a tracer a
// physics package leaves unchanged; with accum it leaves the output alone).
// AK_MASSFIX applies the dry-mass fixer to ln ps in place (dyn_run.f90:626-633
// with the update of run.f90:535-538): column a holds the analysed deviation
// field of log(ps pcorr), scale is dt, and the output s becomes s + dt (t - s) / dt
// with t the analysis plus, at n = 0, s.
enum AnalKind : int { AK_COPY = 0, AK_VRT = 1, AK_DIV = 2, AK_ZERO = 3, AK_MASSFIX = 4 };
struct AnalDesc {
    int kind;
    int a, b;      // analysed columns (field indices): COPY uses a; VRT/DIV use (U~=a, V~=b)
    double scale;  // multiplier of the result
    int inv;       // 1: also multiply by invlap(n) (SHTns S/T instead of vorticity/divergence)
    int accum;     // 1: out += result, 0: out = result
    int c;         // optional third column (-1: none) added as sc * n(n+1) * X_c
    double sc;
};

// Kernel launchers (sht_kernels.hip), all asynchronous on stream s.
// Spectral source element (field f, coefficient k) sits at src[f*sf + k*sk] (complex).
// nfdyn, where given, is a device-resident field count that an earlier kernel of
// the stream wrote: the launch is sized for the static count (its maximum) and
// the workgroups past *nfdyn return at once; the scatter writes zero for an
// output whose source field lies past it (leaves it alone with accum). The
// scatter takes AK_MASSFIX descriptors only with massfix set.
template <typename T>
void launch_synth_gather(const Plan<T> &p, const cplx<T> *src, long sf, long sk, const SynthDesc *desc,
                         int nfo, T *sx, int ldx, double rsphere, hipStream_t s, T *mean = nullptr,
                         const int *nfdyn = nullptr);
template <typename T>
void launch_leg_synth(const Plan<T> &p, const T *sx, int ldx, T *fs, int ldf, int ncol, hipStream_t s,
                      const int *nfdyn = nullptr);
// fmap (with rdx = 1 / a for its x-gradients) maps the grid fields to the
// Legendre fields they read; the mean is indexed by the Legendre field
template <typename T>
void launch_fft_synth(const Plan<T> &p, const T *fs, int ldf, T *grid, long gstride, int nf,
                      const unsigned char *rcos_flag, hipStream_t s, const T *mean = nullptr,
                      const int *slot = nullptr, int fofs = 0, const int *nfdyn = nullptr,
                      const FieldMap *fmap = nullptr, double rdx = 0.0);
template <typename T>
void launch_fft_anal(const Plan<T> &p, const T *grid, long gstride, int nf, const unsigned char *rcos_flag,
                     T *fa, int ldf, hipStream_t s, const int *nfdyn = nullptr);
template <typename T>
void launch_leg_anal(const Plan<T> &p, const T *fa, int ldf, T *sx, int ldx, int ncol, hipStream_t s,
                     const int *nfdyn = nullptr);
template <typename T>
void launch_anal_scatter(const Plan<T> &p, const T *sx, int ldx, const AnalDesc *desc, int nfo,
                         cplx<T> *dst, long df, long dk, double rsphere, hipStream_t s,
                         const int *nfdyn = nullptr, bool massfix = false);

// Workspace for one batched transform pipeline.
template <typename T> struct Work {
    int ncap = 0;          // field capacity
    int ldc = 0;           // row pitch: 2*ncap rounded to COL_PAD, plus COL_SKEW
    T *sx = nullptr;       // nlme x ldc
    T *fs = nullptr;       // nlat x ldc x (T+1) (synthesis Fourier, laid out by fs_index)
    T *fa = nullptr;       // (T+1) x 2 x nhp x ldc (analysis Fourier, zero padded)
    T *grid = nullptr;     // ncap x nlat x nlon
};
template <typename T> int work_alloc(Work<T> &w, const Plan<T> &p, int nfield);
template <typename T> void work_free(Work<T> &w);

}  // namespace gfsdc
