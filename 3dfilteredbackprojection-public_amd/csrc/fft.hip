// fft.hip -- batched row FFTs, the Fourier-space ramp filter and the staged
// projection layout, for gfx950.
//
// The plain row FFTs (FFFT) run out of LDS: one workgroup owns one complex
// row of length n <= 8192, loads it in bit-reversed order, runs log2(n)
// radix-2 butterfly stages against an LDS twiddle table and writes it back.
// The fused filter packs two real detector rows into one complex FFT (the
// gain is real and even, so the real and imaginary parts of the inverse
// carry the two filtered rows independently): for n = 512 .. 2048 one wave
// owns the pair and exchanges data only through its own LDS buffer, for
// larger n a workgroup runs radix-8 Stockham passes through LDS.
#include <mutex>

#include "ctr_internal.h"

namespace ctr {

constexpr int kMaxLog2 = 13;                 // 8192-point rows
constexpr int kMaxN = 1 << kMaxLog2;
constexpr int kFftThreads = 256;

// exp(-2 pi i k / 8192), k < 4096, rounded from FP64.  Filled once per
// device by init_twiddles; a row of length n uses every (8192/n)-th entry.
__device__ float2 g_twiddle[kMaxN / 2];

__global__ void init_twiddles_kernel()
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < kMaxN / 2) {
        double s, c;
        sincospi(-(double)k / (double)(kMaxN / 2), &s, &c);
        g_twiddle[k] = make_float2((float)c, (float)s);
    }
}

// Filter gains of the unshifted spectrum, per filter: entry n + k holds the
// gain of index k of an n-point row (n a power of two <= 8192), the same
// value unshifted_gain computes; filled once per device next to the
// twiddles so that the row filters read a table instead of evaluating the
// FP64 sine of the Shepp-Logan / cosine gains per sample.
__device__ float g_gain[3][2 * kMaxN];

__device__ inline float filter_g(int filter, double x);
__device__ inline float unshifted_gain(int filter, int k, int n);

__global__ void init_gains_kernel()
{
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;  // over 3 * 2 * kMaxN
    if (idx >= 3 * 2 * kMaxN) return;
    const int f = idx / (2 * kMaxN), e = idx % (2 * kMaxN);
    if (e == 0) {
        g_gain[f][0] = 0.0f;
        return;
    }
    const int n = 1 << (31 - __clz(e));  // largest power of two <= e
    g_gain[f][e] = unshifted_gain(f, e - n, n);
}

static std::mutex g_tw_lock;
static bool g_tw_ready[64] = {false};

static int ensure_twiddles()
{
    int dev = 0;
    CTR_TRY_HIP(hipGetDevice(&dev));
    CTR_REQUIRE(dev >= 0 && dev < 64, CTR_ERR_INVALID_ARGUMENT,
                "device index %d out of range", dev);
    std::lock_guard<std::mutex> g(g_tw_lock);
    if (!g_tw_ready[dev]) {
        init_twiddles_kernel<<<kMaxN / 2 / 256, 256, 0, 0>>>();
        CTR_CHECK_LAUNCH();
        init_gains_kernel<<<3 * 2 * kMaxN / 256, 256, 0, 0>>>();
        CTR_CHECK_LAUNCH();
        CTR_TRY_HIP(hipDeviceSynchronize());
        g_tw_ready[dev] = true;
    }
    return CTR_OK;
}

__device__ inline float2 cmul(float2 a, float2 b)
{
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

template <int LOG2N>
__device__ inline unsigned bitrev(unsigned e)
{
    return LOG2N == 0 ? 0u : (__brev(e) >> (32 - LOG2N));
}

// Loads the n/2 twiddles of an n-point transform into LDS.
template <int LOG2N>
__device__ inline void load_twiddles(float2 *T)
{
    constexpr int N = 1 << LOG2N;
    constexpr int stride = kMaxN / N;
    for (int k = threadIdx.x; k < N / 2; k += blockDim.x) T[k] = g_twiddle[k * stride];
}

// In-place radix-2 DIT stages over bit-reversed data in LDS.  INV selects
// the conjugate twiddles (conj -> FFT -> conj of the reference equals this
// exactly).  Ends with a barrier.
template <int LOG2N, bool INV>
__device__ inline void lds_stages(float2 *L, const float2 *T)
{
    constexpr int N = 1 << LOG2N;
#pragma unroll 1
    for (int s = 0; s < LOG2N; ++s) {
        const int half = 1 << s;
        const int tshift = LOG2N - 1 - s;  // twiddle index = pos << tshift
        for (int b = threadIdx.x; b < N / 2; b += blockDim.x) {
            int pos = b & (half - 1);
            int i0 = ((b >> s) << (s + 1)) + pos;
            int i1 = i0 + half;
            float2 w = T[pos << tshift];
            if (INV) w.y = -w.y;
            float2 x0 = L[i0];
            float2 t = cmul(L[i1], w);
            L[i0] = make_float2(x0.x + t.x, x0.y + t.y);
            L[i1] = make_float2(x0.x - t.x, x0.y - t.y);
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Complex row FFT (FFFT::fft_radix2 / inverse_fft_radix2)
// ---------------------------------------------------------------------------
template <int LOG2N, bool INV>
__global__ __launch_bounds__(kFftThreads) void fft_rows_kernel(float2 *data, int rows)
{
    constexpr int N = 1 << LOG2N;
    __shared__ float2 L[N];
    __shared__ float2 T[N / 2 > 0 ? N / 2 : 1];
    const size_t base = (size_t)blockIdx.x * N;
    load_twiddles<LOG2N>(T);
    for (int e = threadIdx.x; e < N; e += blockDim.x) L[bitrev<LOG2N>(e)] = data[base + e];
    __syncthreads();
    lds_stages<LOG2N, INV>(L, T);
    const float inv_n = 1.0f / (float)N;
    for (int e = threadIdx.x; e < N; e += blockDim.x) {
        float2 v = L[e];
        if (INV) v = make_float2(v.x * inv_n, v.y * inv_n);
        data[base + e] = v;
    }
}

template <bool INV>
static int launch_fft_rows(hipStream_t s, float2 *data, int log2n, int rows)
{
#define CTR_FFT_CASE(L)                                                          \
    case L:                                                                      \
        fft_rows_kernel<L, INV><<<rows, kFftThreads, 0, s>>>(data, rows);        \
        break;
    switch (log2n) {
        CTR_FFT_CASE(0) CTR_FFT_CASE(1) CTR_FFT_CASE(2) CTR_FFT_CASE(3)
        CTR_FFT_CASE(4) CTR_FFT_CASE(5) CTR_FFT_CASE(6) CTR_FFT_CASE(7)
        CTR_FFT_CASE(8) CTR_FFT_CASE(9) CTR_FFT_CASE(10) CTR_FFT_CASE(11)
        CTR_FFT_CASE(12) CTR_FFT_CASE(13)
    default:
        return fail(CTR_ERR_NOT_POWER_OF_TWO, "FFT length 2^%d unsupported", log2n);
    }
#undef CTR_FFT_CASE
    CTR_CHECK_LAUNCH();
    return CTR_OK;
}

// ---------------------------------------------------------------------------
// Filter gain
// ---------------------------------------------------------------------------
// g(x) of FastFourierTransformImage.cpp:35-37 with true pi.
__device__ inline float filter_g(int filter, double x)
{
    x = fabs(x);
    if (filter == CTR_RAM_LAK) return (float)x;
    if (filter == CTR_SHEPP_LOGAN) return (float)sin(x);
    return (float)(fabs(sinpi(x)) / M_PI);
}

// Gain of the UNSHIFTED spectrum index k of an n-point row (n a power of
// two): fft_shift moves k to i = (k + n/2) mod n, whose gain is
// n * g((float)i/n - 0.5f) (FastFourierTransformImage.cpp:45), so the shift,
// multiply and inverse shift of the reference collapse into one multiply.
__device__ inline float unshifted_gain(int filter, int k, int n)
{
    int i = (k + n / 2) & (n - 1);
    float x = (float)i / (float)n - 0.5f;
    return (float)n * filter_g(filter, (double)x);
}

__global__ void apply_filter_kernel(float2 *spec, int n, long long total, int filter)
{
    long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= total) return;
    int i = (int)(idx % n);
    // FastFourierTransformImage.cpp:45: x = (float)i/n - 0.5f, times n
    float x = (float)i / (float)n - 0.5f;
    float g = (float)n * filter_g(filter, (double)x);
    float2 v = spec[idx];
    spec[idx] = make_float2(v.x * g, v.y * g);
}

// ---------------------------------------------------------------------------
// Fused pad -> FFT -> gain -> IFFT -> crop over row pairs
// ---------------------------------------------------------------------------
struct FilterJob {
    const float *in;
    int width;          // samples per row (nu)
    int rows;           // total rows
    int filter;
    // optional FDK pre-weight
    int fdk;
    int layout;         // CTR_LAYOUT_* of the input rows
    int nv, n_views;
    double sdd2, delta;
    // output: plain rows (staged == nullptr) or fp32 staged layout
    float *out;
    float *staged;
    int pitch_u, pitch_v;
    // optional per-row max |value| of the output (indexed like the input
    // rows), reduced per view afterwards
    float *row_max;
};

// Stores the row pair's max |value|: wave reduction, then one LDS pass over
// the (at most 16) waves of the workgroup; no atomics.
__device__ inline void note_row_max(const FilterJob &J, float ma, float mb, int r0, bool has1)
{
    __shared__ float part[2][16];
    for (int off = 32; off > 0; off >>= 1) {
        ma = fmaxf(ma, __shfl_xor(ma, off));
        mb = fmaxf(mb, __shfl_xor(mb, off));
    }
    const int wave = threadIdx.x >> 6, waves = (blockDim.x + 63) >> 6;
    if ((threadIdx.x & 63) == 0) {
        part[0][wave] = ma;
        part[1][wave] = mb;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < waves; ++w) {
            ma = fmaxf(ma, part[0][w]);
            mb = fmaxf(mb, part[1][w]);
        }
        J.row_max[r0] = ma;
        if (has1) J.row_max[r0 + 1] = mb;
    }
}

// view_scale[n] = max over the rows of view n of row_max (1 when 0).
__global__ __launch_bounds__(256) void view_max_from_rows_kernel(const float *row_max, int nv,
                                                                 float *view_scale)
{
    // row maxima in projection order (row = view * nv + iv), as the filter
    // numbers its rows in either input layout
    const int n = blockIdx.x;
    float m = 0.0f;
    for (int iv = threadIdx.x; iv < nv; iv += blockDim.x) m = fmaxf(m, row_max[(size_t)n * nv + iv]);
    for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
    __shared__ float part[4];
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        m = fmaxf(fmaxf(part[0], part[1]), fmaxf(part[2], part[3]));
        view_scale[n] = m > 0.0f ? m : 1.0f;
    }
}

// FDK cosine weight SDD / sqrt(SDD^2 + u^2 + v^2) of detector column e for a
// row with vv = SDD^2 + v^2, in FP64: the hardware reciprocal square root
// refined by one Newton step y (3 - x y^2) / 2 (about 2^-44 relative), a
// handful of FP64 operations instead of a library sqrt and divide.
__device__ inline float fdk_weight_at(double sdd, double vv, double cu, double delta, int e)
{
    const double u = ((double)e - cu) * delta;
    const double x = fma(u, u, vv);
    double y = __builtin_amdgcn_rsq(x);
    y *= fma(-0.5 * x * y, y, 1.5);
    return (float)(sdd * y);
}

// Rows are numbered in projection order, r = view * nv + iv, whatever the
// input layout, and transformed in pairs (2p, 2p + 1): the same two rows
// share a complex FFT for projections and for sinograms, so both layouts
// give the same staged values bit for bit.  row_coords gives (view, iv) of
// row r, input_row its row in the input.
__device__ inline void row_coords(const FilterJob &J, int r, int &view, int &iv)
{
    iv = r % J.nv;
    view = r / J.nv;
}

__device__ inline size_t input_row(const FilterJob &J, int r)
{
    if (J.layout == CTR_LAYOUT_PROJECTIONS) return (size_t)r;
    return (size_t)(r % J.nv) * J.n_views + r / J.nv;
}

template <int LOG2N>
__global__ __launch_bounds__(kFftThreads) void filter_pairs_kernel(FilterJob J)
{
    constexpr int N = 1 << LOG2N;
    __shared__ float2 L[N];
    __shared__ float2 T[N / 2 > 0 ? N / 2 : 1];
    constexpr int PER = (N + kFftThreads - 1) / kFftThreads;

    const int r0 = 2 * blockIdx.x, r1 = r0 + 1;
    const bool has1 = r1 < J.rows;
    int view0 = 0, iv0 = 0, view1 = 0, iv1 = 0;
    row_coords(J, r0, view0, iv0);
    if (has1) row_coords(J, r1, view1, iv1);
    double vv0 = 0.0, vv1 = 0.0, cu = 0.5 * J.width - 0.5;
    if (J.fdk) {
        double v0 = (iv0 + 0.5 - 0.5 * J.nv) * J.delta;
        double v1 = (iv1 + 0.5 - 0.5 * J.nv) * J.delta;
        vv0 = J.sdd2 + v0 * v0;
        vv1 = J.sdd2 + v1 * v1;
    }
    const float *a_row = J.in + input_row(J, r0) * J.width;
    const float *b_row = J.in + input_row(J, has1 ? r1 : r0) * J.width;
    const double sdd = J.fdk ? sqrt(J.sdd2) : 0.0;

    load_twiddles<LOG2N>(T);
    for (int e = threadIdx.x; e < N; e += blockDim.x) {
        float a = 0.f, b = 0.f;
        if (e < J.width) {
            a = a_row[e];
            if (has1) b = b_row[e];
            if (J.fdk) {
                a *= fdk_weight_at(sdd, vv0, cu, J.delta, e);
                b *= fdk_weight_at(sdd, vv1, cu, J.delta, e);
            }
        }
        L[bitrev<LOG2N>(e)] = make_float2(a, b);
    }
    __syncthreads();
    lds_stages<LOG2N, false>(L, T);

    // gain, then bit-reverse again for the inverse transform
    float2 reg[PER];
#pragma unroll
    for (int p = 0; p < PER; ++p) {
        int k = threadIdx.x + p * kFftThreads;
        if (k < N) {
            float g = g_gain[J.filter][N + k];
            float2 v = L[k];
            reg[p] = make_float2(v.x * g, v.y * g);
        }
    }
    __syncthreads();
#pragma unroll
    for (int p = 0; p < PER; ++p) {
        int k = threadIdx.x + p * kFftThreads;
        if (k < N) L[bitrev<LOG2N>(k)] = reg[p];
    }
    __syncthreads();
    lds_stages<LOG2N, true>(L, T);

    const float inv_n = 1.0f / (float)N;
    if (J.staged == nullptr) {
        float *oa = J.out + (size_t)r0 * J.width;
        float *ob = J.out + (size_t)r1 * J.width;
        for (int e = threadIdx.x; e < J.width; e += blockDim.x) {
            float2 v = L[e];
            oa[e] = v.x * inv_n;
            if (has1) ob[e] = v.y * inv_n;
        }
    } else {
        size_t vstride = (size_t)J.pitch_u * J.pitch_v;
        float *oa = J.staged + (size_t)view0 * vstride + (size_t)(iv0 + 1) * J.pitch_u + 1;
        float *ob = J.staged + (size_t)view1 * vstride + (size_t)(iv1 + 1) * J.pitch_u + 1;
        float ma = 0.0f, mb = 0.0f;
        for (int e = threadIdx.x; e < J.width; e += blockDim.x) {
            float2 v = L[e];
            const float a = v.x * inv_n, b = v.y * inv_n;
            oa[e] = a;
            if (has1) ob[e] = b;
            ma = fmaxf(ma, fabsf(a));
            mb = fmaxf(mb, fabsf(b));
        }
        if (J.row_max) note_row_max(J, ma, mb, r0, has1);
    }
}

// ---------------------------------------------------------------------------
// Stockham radix-8 row filter (n >= 512)
// ---------------------------------------------------------------------------
// n/8 threads per row pair.  Every thread keeps 8 complex values in
// registers through a pass and exchanges them through LDS between passes
// (read, barrier, radix-R butterflies, write, barrier): radix-8 passes plus
// a final radix-2 or radix-4 pass when log2 n is not a multiple of 3, so a
// 2048-point transform takes 4 LDS round trips instead of 11.  The first
// forward pass reads the detector rows straight from memory, the gain is
// applied as the inverse transform reads the spectrum, and the last inverse
// pass writes the cropped rows straight to memory.  LDS indices carry one
// float2 of padding per 32 against bank conflicts of the stride-8 writes.

__device__ inline int lpad(int i) { return i + (i >> 5); }

template <bool INV>
__device__ inline float2 times_minus_i(float2 a)  // a * (-i) forward, a * (+i) inverse
{
    return INV ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}

__device__ inline float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ inline float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }

template <int R, bool INV>
struct SmallFFT;

template <bool INV>
struct SmallFFT<2, INV> {
    __device__ static inline void run(float2 *x)
    {
        float2 a = x[0], b = x[1];
        x[0] = cadd(a, b);
        x[1] = csub(a, b);
    }
};

template <bool INV>
struct SmallFFT<4, INV> {
    __device__ static inline void run(float2 *x)
    {
        float2 y0 = cadd(x[0], x[2]), y1 = csub(x[0], x[2]);
        float2 y2 = cadd(x[1], x[3]), y3 = times_minus_i<INV>(csub(x[1], x[3]));
        x[0] = cadd(y0, y2);
        x[2] = csub(y0, y2);
        x[1] = cadd(y1, y3);
        x[3] = csub(y1, y3);
    }
};

template <bool INV>
struct SmallFFT<8, INV> {
    __device__ static inline void run(float2 *x)
    {
        const float h = 0.70710678118654752440f;  // 1/sqrt(2)
        float2 e[4] = {x[0], x[2], x[4], x[6]};
        float2 o[4] = {x[1], x[3], x[5], x[7]};
        SmallFFT<4, INV>::run(e);
        SmallFFT<4, INV>::run(o);
        // o_k *= exp(-+2 pi i k / 8)
        float2 a = o[1];
        o[1] = INV ? make_float2((a.x - a.y) * h, (a.x + a.y) * h)
                   : make_float2((a.x + a.y) * h, (a.y - a.x) * h);
        o[2] = times_minus_i<INV>(o[2]);
        a = o[3];
        o[3] = INV ? make_float2(-(a.x + a.y) * h, (a.x - a.y) * h)
                   : make_float2((a.y - a.x) * h, -(a.x + a.y) * h);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            x[k] = cadd(e[k], o[k]);
            x[k + 4] = csub(e[k], o[k]);
        }
    }
};

// Radix-R butterflies of one Stockham pass with span Ns on the 8 / R
// butterflies this thread owns (v[b * R + r]).
template <int N, int R, bool INV>
__device__ inline void stockham_butterflies(float2 (&v)[8], int Ns, const float2 *Tw, int tid)
{
    constexpr int T = N / 8, B = 8 / R;
#pragma unroll
    for (int b = 0; b < B; ++b) {
        const int j = tid + b * T;
        const int jm = j & (Ns - 1);
        if (Ns > 1) {
            const int step = (N / R) / Ns;
#pragma unroll
            for (int r = 1; r < R; ++r) {
                float2 w = Tw[jm * r * step];
                if (INV) w.y = -w.y;
                v[b * R + r] = cmul(v[b * R + r], w);
            }
        }
        SmallFFT<R, INV>::run(&v[b * R]);
    }
}

template <int N, int R>
__device__ inline void stockham_read(float2 (&v)[8], const float2 *L, int tid)
{
    constexpr int T = N / 8, B = 8 / R;
#pragma unroll
    for (int b = 0; b < B; ++b)
#pragma unroll
        for (int r = 0; r < R; ++r) v[b * R + r] = L[lpad(tid + b * T + r * (N / R))];
}

template <int N, int R>
__device__ inline void stockham_write(const float2 (&v)[8], float2 *L, int Ns, int tid)
{
    constexpr int T = N / 8, B = 8 / R;
#pragma unroll
    for (int b = 0; b < B; ++b) {
        const int j = tid + b * T;
        const int jm = j & (Ns - 1);
        const int base = (j - jm) * R + jm;
#pragma unroll
        for (int r = 0; r < R; ++r) L[lpad(base + r * Ns)] = v[b * R + r];
    }
}

// One full in-LDS pass (read, barrier, butterflies, write, barrier).
template <int N, int R, bool INV>
__device__ inline void stockham_pass(float2 *L, const float2 *Tw, int Ns, int tid)
{
    float2 v[8];
    stockham_read<N, R>(v, L, tid);
    __syncthreads();
    stockham_butterflies<N, R, INV>(v, Ns, Tw, tid);
    stockham_write<N, R>(v, L, Ns, tid);
    __syncthreads();
}

template <int LOG2N>
__global__ __launch_bounds__((1 << LOG2N) / 8) void filter_stockham_kernel(FilterJob J)
{
    constexpr int N = 1 << LOG2N, T = N / 8;
    constexpr int P8 = LOG2N / 3;             // radix-8 passes
    constexpr int RL = 1 << (LOG2N % 3);      // final radix (1: none)
    static_assert(P8 >= 2, "the Stockham filter needs n >= 64");
    __shared__ float2 L[N + N / 32];
    __shared__ float2 Tw[N];
    const int tid = threadIdx.x;

    // full-circle twiddle table exp(-2 pi i k / N)
    for (int k = tid; k < N / 2; k += T) {
        float2 w = g_twiddle[k * (kMaxN / N)];
        Tw[k] = w;
        Tw[k + N / 2] = make_float2(-w.x, -w.y);
    }

    const int r0 = 2 * blockIdx.x, r1 = r0 + 1;
    const bool has1 = r1 < J.rows;
    int view0 = 0, iv0 = 0, view1 = 0, iv1 = 0;
    row_coords(J, r0, view0, iv0);
    if (has1) row_coords(J, r1, view1, iv1);
    double vv0 = 0.0, vv1 = 0.0, sdd = 0.0;
    const double cu = 0.5 * J.width - 0.5;
    if (J.fdk) {
        double v0 = (iv0 + 0.5 - 0.5 * J.nv) * J.delta;
        double v1 = (iv1 + 0.5 - 0.5 * J.nv) * J.delta;
        vv0 = J.sdd2 + v0 * v0;
        vv1 = J.sdd2 + v1 * v1;
        sdd = sqrt(J.sdd2);
    }
    const float *a_row = J.in + input_row(J, r0) * J.width;
    const float *b_row = J.in + input_row(J, has1 ? r1 : r0) * J.width;

    // forward pass 0 (radix 8, span 1) straight from the detector rows,
    // through buffer resources sized to the rows (columns past the row and
    // a missing second row read zero): the eight loads go out together
    const __amdgpu_buffer_rsrc_t ra = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<float *>(a_row), (short)0, J.width * 4, 0x00020000);
    const __amdgpu_buffer_rsrc_t rb = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<float *>(has1 ? b_row : a_row), (short)0, has1 ? J.width * 4 : 0, 0x00020000);
    float2 v[8];
#pragma unroll
    for (int r = 0; r < 8; ++r) {
        const int off = (tid + r * T) * 4;
        v[r] = make_float2(
            __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(ra, off, 0, 0)),
            __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rb, off, 0, 0)));
    }
    if (J.fdk) {
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            const int e = tid + r * T;
            if (e < J.width) {
                v[r].x *= fdk_weight_at(sdd, vv0, cu, J.delta, e);
                v[r].y *= fdk_weight_at(sdd, vv1, cu, J.delta, e);
            }
        }
    }
    SmallFFT<8, false>::run(v);
    stockham_write<N, 8>(v, L, 1, tid);
    __syncthreads();
    int Ns = 8;
#pragma unroll 1
    for (int p = 1; p < P8; ++p, Ns *= 8) stockham_pass<N, 8, false>(L, Tw, Ns, tid);
    if (RL > 1) stockham_pass<N, (RL > 1 ? RL : 2), false>(L, Tw, Ns, tid);

    // inverse pass 0 (radix 8, span 1) with the gain applied on read
#pragma unroll
    for (int r = 0; r < 8; ++r) {
        const int e = tid + r * T;
        const float g = g_gain[J.filter][N + e];
        const float2 x = L[lpad(e)];
        v[r] = make_float2(x.x * g, x.y * g);
    }
    __syncthreads();
    SmallFFT<8, true>::run(v);
    stockham_write<N, 8>(v, L, 1, tid);
    __syncthreads();
    Ns = 8;
    constexpr int mid_end = RL > 1 ? P8 : P8 - 1;  // in-LDS radix-8 passes
#pragma unroll 1
    for (int p = 1; p < mid_end; ++p, Ns *= 8) stockham_pass<N, 8, true>(L, Tw, Ns, tid);

    // last inverse pass: span N/R, so element index = j + r * N/R
    constexpr int RF = RL > 1 ? RL : 8;
    stockham_read<N, RF>(v, L, tid);
    stockham_butterflies<N, RF, true>(v, Ns, Tw, tid);
    const float inv_n = 1.0f / (float)N;
    float *oa, *ob;
    if (J.staged == nullptr) {
        oa = J.out + (size_t)r0 * J.width;
        ob = J.out + (size_t)r1 * J.width;
    } else {
        size_t vstride = (size_t)J.pitch_u * J.pitch_v;
        oa = J.staged + (size_t)view0 * vstride + (size_t)(iv0 + 1) * J.pitch_u + 1;
        ob = J.staged + (size_t)view1 * vstride + (size_t)(iv1 + 1) * J.pitch_u + 1;
    }
    float ma = 0.0f, mb = 0.0f;
#pragma unroll
    for (int b = 0; b < 8 / RF; ++b)
#pragma unroll
        for (int r = 0; r < RF; ++r) {
            const int e = tid + b * T + r * (N / RF);
            if (e < J.width) {
                const float x = v[b * RF + r].x * inv_n, y = v[b * RF + r].y * inv_n;
                oa[e] = x;
                if (has1) ob[e] = y;
                ma = fmaxf(ma, fabsf(x));
                mb = fmaxf(mb, fabsf(y));
            }
        }
    if (J.row_max) note_row_max(J, ma, mb, r0, has1);
}

// ---------------------------------------------------------------------------
// One-wave row filter (n = 512, 1024, 2048)
// ---------------------------------------------------------------------------
// A wave owns a pair of rows (one complex n-point FFT, n = 64 M) and keeps
// it in registers: lane l holds x[l + 64 m], m < M.  The forward transform
// is the four-step factorisation n = M x 64 with the 64-point part split
// 8 x 8 (l = l1 + 8 l2, p = p2 + 8 p1, output k = q + M (p2 + 8 p1)):
//   1. an M-point DFT over m in registers, times W_n^(l q);
//   2. exchange through the wave's own LDS buffer, then 8-point DFTs over
//      l2 (lane l1 + 8 qb owns q = (M/8) qb + j, j < M/8), times
//      W_64^(l1 p2);
//   3. exchange, then 8-point DFTs over l1 (lanes p2 + 8 qb).
// The DFT matrix is symmetric, so the inverse runs the same steps in
// reverse order with conjugate twiddles and ends in the input layout: no
// reordering pass.  The exchanges are wave-private (no workgroup barrier;
// LDS operations of one wave complete in order) and run one j group per
// round, which keeps the buffer at 8 rows of 72 complex values.  The rows
// are read through buffer resources, so all loads of a row pair are issued
// together.  Workgroups are persistent: the twiddle tables are built once
// in LDS and each wave loops over row pairs.

// One-wave filter for n = 512 .. 2048 (1) or the Stockham kernel (0); j
// groups per exchange step; minimum waves per SIMD the compiler must keep
// registers for.  Measured on 512^3 x 720 (FFT 2048, fp32 staging, ms):
// Stockham 2.81, one-wave 1.94 (2 waves per SIMD) / 1.86 (3, chosen);
// 256^3 x 360 (FFT 1024) 0.337 / 0.214; two j groups per exchange step 0.266.
#ifndef CTR_WAVE_FILTER
#define CTR_WAVE_FILTER 1
#endif
#ifndef CTR_WAVE_FILTER_JH
#define CTR_WAVE_FILTER_JH 1
#endif
#ifndef CTR_WAVE_FILTER_MINW
#define CTR_WAVE_FILTER_MINW 3
#endif

constexpr double cexpr_pi = 3.141592653589793238462643383279502884;

// cos / sin of 2 pi e / n at compile time (Taylor series after folding the
// angle into [-pi, pi]).
constexpr double cexpr_cos_sin(int e, int n, bool want_sin)
{
    e %= n;
    if (e < 0) e += n;
    double x = 2.0 * cexpr_pi * (double)e / (double)n;
    if (x > cexpr_pi) x -= 2.0 * cexpr_pi;
    double term = want_sin ? x : 1.0, sum = term;
    for (int k = 1; k < 30; ++k) {
        const int a = want_sin ? 2 * k : 2 * k - 1, b = want_sin ? 2 * k + 1 : 2 * k;
        term *= -x * x / ((double)a * (double)b);
        sum += term;
    }
    return sum;
}

// x * W_n^e (forward) or x * conj(W_n^e), W_n = exp(-2 pi i / n), e known
// at compile time.
template <int N, int E, bool INV>
__device__ inline float2 ctwiddle(float2 x)
{
    if constexpr (E % N == 0) {
        return x;
    } else {
        constexpr float c = (float)cexpr_cos_sin(E, N, false);
        constexpr float s = INV ? (float)cexpr_cos_sin(E, N, true)
                                : -(float)cexpr_cos_sin(E, N, true);
        return make_float2(x.x * c - x.y * s, x.x * s + x.y * c);
    }
}

// M-point DFT (M = 8, 16, 32) of v in registers, natural order in and out:
// M = 8 R, X[k8 + 8 kr] = sum_r W_R^(r kr) W_M^(r k8) sum_s W_8^(s k8) v[r + R s].
template <int M, bool INV, int R>
struct LaneDftStage {
    template <int r, int k8>
    __device__ static inline void tw(float2 (&a)[R][8])
    {
        if constexpr (k8 < 8) {
            a[r][k8] = ctwiddle<M, r * k8, INV>(a[r][k8]);
            tw<r, k8 + 1>(a);
        }
    }
    template <int r>
    __device__ static inline void tw_all(float2 (&a)[R][8])
    {
        if constexpr (r < R) {
            tw<r, 1>(a);
            tw_all<r + 1>(a);
        }
    }
};

template <int M, bool INV>
__device__ inline void lane_dft(float2 (&v)[M])
{
    if constexpr (M == 8) {
        SmallFFT<8, INV>::run(v);
    } else {
        constexpr int R = M / 8;
        float2 a[R][8];
        // scheduling barriers between the sub-DFTs keep the compiler from
        // interleaving all of them (which needs twice the registers)
#pragma unroll
        for (int r = 0; r < R; ++r) {
#pragma unroll
            for (int s = 0; s < 8; ++s) a[r][s] = v[r + R * s];
            SmallFFT<8, INV>::run(a[r]);
            __builtin_amdgcn_sched_barrier(0);
        }
        LaneDftStage<M, INV, R>::template tw_all<1>(a);
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int k8 = 0; k8 < 8; ++k8) {
            float2 z[R];
#pragma unroll
            for (int r = 0; r < R; ++r) z[r] = a[r][k8];
            SmallFFT<R, INV>::run(z);
#pragma unroll
            for (int kr = 0; kr < R; ++kr) v[k8 + 8 * kr] = z[kr];
            if (k8 & 1) __builtin_amdgcn_sched_barrier(0);
        }
    }
}

// Orders the LDS accesses of one wave across lanes (the hardware executes a
// wave's LDS instructions in order; this keeps the compiler from moving
// them across the exchange points).
__device__ inline void wave_lds_fence()
{
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_sched_barrier(0);
}

template <int LOG2N>
struct WaveFft {
    static constexpr int N = 1 << LOG2N, M = N / 64, J4 = M / 8;
    static constexpr int JH = J4 >= 2 ? CTR_WAVE_FILTER_JH : 1, H = J4 / JH;
    // row strides of the exchange buffer (complex elements), chosen so that
    // the strided read of each exchange hits 16 distinct bank pairs
    static constexpr int R1 = JH == 2 ? 68 : 72, R2 = 65;
    static constexpr int kBuf = 8 * JH * (R1 > R2 ? R1 : R2);
    static_assert(M == 8 || M == 16 || M == 32, "n = 512, 1024 or 2048");

    // forward: input layout -> output layout (slot p1 J4 + j holds
    // X[qb J4 + j + M (p2 + 8 p1)], lane = p2 + 8 qb)
    __device__ static inline void forward(float2 (&v)[M], float2 *B, const float2 *TwM,
                                          const float2 *Tw64, int lane)
    {
        lane_dft<M, false>(v);
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int q = 1; q < M; ++q) {
            v[q] = cmul(v[q], TwM[q * 64 + lane]);
            if ((q & 7) == 7) __builtin_amdgcn_sched_barrier(0);
        }
        const int lo = lane & 7, hi = lane >> 3;
#pragma unroll
        for (int h = 0; h < H; ++h) {
            wave_lds_fence();
#pragma unroll
            for (int qb = 0; qb < 8; ++qb)
#pragma unroll
                for (int jj = 0; jj < JH; ++jj)
                    B[(qb * JH + jj) * R1 + lane] = v[qb * J4 + h * JH + jj];
            wave_lds_fence();
            float2 y[JH][8];
#pragma unroll
            for (int jj = 0; jj < JH; ++jj)
#pragma unroll
                for (int l2 = 0; l2 < 8; ++l2) y[jj][l2] = B[(hi * JH + jj) * R1 + lo + 8 * l2];
            wave_lds_fence();
#pragma unroll
            for (int jj = 0; jj < JH; ++jj) {
                SmallFFT<8, false>::run(y[jj]);
#pragma unroll
                for (int p2 = 1; p2 < 8; ++p2) y[jj][p2] = cmul(y[jj][p2], Tw64[p2 * 8 + lo]);
#pragma unroll
                for (int p2 = 0; p2 < 8; ++p2) B[(jj * 8 + p2) * R2 + lane] = y[jj][p2];
            }
            wave_lds_fence();
#pragma unroll
            for (int jj = 0; jj < JH; ++jj) {
                float2 z[8];
#pragma unroll
                for (int l1 = 0; l1 < 8; ++l1) z[l1] = B[(jj * 8 + lo) * R2 + l1 + 8 * hi];
                SmallFFT<8, false>::run(z);
#pragma unroll
                for (int p1 = 0; p1 < 8; ++p1) v[p1 * J4 + h * JH + jj] = z[p1];
            }
        }
    }

    // inverse (unnormalised, conjugate twiddles): output layout -> input
    // layout, the forward steps transposed
    __device__ static inline void inverse(float2 (&v)[M], float2 *B, const float2 *TwM,
                                          const float2 *Tw64, int lane)
    {
        const int lo = lane & 7, hi = lane >> 3;
#pragma unroll
        for (int h = 0; h < H; ++h) {
            wave_lds_fence();
#pragma unroll
            for (int jj = 0; jj < JH; ++jj) {
                float2 z[8];
#pragma unroll
                for (int p1 = 0; p1 < 8; ++p1) z[p1] = v[p1 * J4 + h * JH + jj];
                SmallFFT<8, true>::run(z);
#pragma unroll
                for (int l1 = 1; l1 < 8; ++l1) {
                    float2 w = Tw64[lo * 8 + l1];
                    w.y = -w.y;
                    z[l1] = cmul(z[l1], w);
                }
#pragma unroll
                for (int l1 = 0; l1 < 8; ++l1) B[(jj * 8 + lo) * R2 + l1 + 8 * hi] = z[l1];
            }
            wave_lds_fence();
            float2 y[JH][8];
#pragma unroll
            for (int jj = 0; jj < JH; ++jj)
#pragma unroll
                for (int p2 = 0; p2 < 8; ++p2) y[jj][p2] = B[(jj * 8 + p2) * R2 + lane];
            wave_lds_fence();
#pragma unroll
            for (int jj = 0; jj < JH; ++jj) {
                SmallFFT<8, true>::run(y[jj]);
                const int q = hi * J4 + h * JH + jj;
#pragma unroll
                for (int l2 = 0; l2 < 8; ++l2) {
                    if (q != 0 || l2 != 0) {
                        float2 w = TwM[q * 64 + lo + 8 * l2];
                        w.y = -w.y;
                        y[jj][l2] = cmul(y[jj][l2], w);
                    }
                    B[(hi * JH + jj) * R1 + lo + 8 * l2] = y[jj][l2];
                }
            }
            wave_lds_fence();
#pragma unroll
            for (int qb = 0; qb < 8; ++qb)
#pragma unroll
                for (int jj = 0; jj < JH; ++jj)
                    v[qb * J4 + h * JH + jj] = B[(qb * JH + jj) * R1 + lane];
        }
        lane_dft<M, true>(v);
    }
};

constexpr int kWaveFftWaves = 4;  // waves (row pairs in flight) per workgroup

template <int LOG2N>
__global__ __launch_bounds__(64 * kWaveFftWaves, CTR_WAVE_FILTER_MINW) void filter_wave_kernel(FilterJob J)
{
    using F = WaveFft<LOG2N>;
    constexpr int N = F::N, M = F::M, J4 = F::J4;
    __shared__ float2 TwM[M * 64];   // W_n^(l q), q < M, l < 64
    __shared__ float2 Tw64[64];      // W_64^(l1 p2) at p2 * 8 + l1
    __shared__ float2 Xch[kWaveFftWaves][F::kBuf];
    const int tid = threadIdx.x, lane = tid & 63;
    // wave-uniform in a scalar register: the row pair, its row coordinates
    // and row addresses then stay scalar
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    constexpr int S = kMaxN / N;
    auto twiddle = [](int e) {
        const float2 w = g_twiddle[(e & (N / 2 - 1)) * S];
        return e < N / 2 ? w : make_float2(-w.x, -w.y);
    };
    for (int i = tid; i < M * 64; i += 64 * kWaveFftWaves) TwM[i] = twiddle((i >> 6) * (i & 63));
    if (tid < 64) Tw64[tid] = twiddle((N / 64) * (tid >> 3) * (tid & 7));
    __syncthreads();

    float2 *B = Xch[wave];
    const int pairs = (J.rows + 1) / 2;
    const double cu = 0.5 * J.width - 0.5;
    const double sdd = J.fdk ? sqrt(J.sdd2) : 0.0;
    const float inv_n = 1.0f / (float)N;
#pragma unroll 1
    for (int rp = blockIdx.x * kWaveFftWaves + wave; rp < pairs;
         rp += gridDim.x * kWaveFftWaves) {
        // the lane index made opaque per row pair: every lane-derived column,
        // weight and gain address is recomputed in the loop rather than
        // hoisted out of it, where the 32 per-column values would stay live
        int ln = lane;
        asm volatile("" : "+v"(ln));
        const int lo = ln & 7, hi = ln >> 3;
        const int r0 = 2 * rp, r1 = r0 + 1;
        const bool has1 = r1 < J.rows;
        const int r1c = has1 ? r1 : r0;  // row coordinates of a valid row
        int view0, iv0, view1, iv1;
        row_coords(J, r0, view0, iv0);
        row_coords(J, r1c, view1, iv1);
        double vv0 = 0.0, vv1 = 0.0;
        if (J.fdk) {
            const double v0 = (iv0 + 0.5 - 0.5 * J.nv) * J.delta;
            const double v1 = (iv1 + 0.5 - 0.5 * J.nv) * J.delta;
            vv0 = J.sdd2 + v0 * v0;
            vv1 = J.sdd2 + v1 * v1;
        }
        // both rows through buffer resources sized to the row: columns past
        // the row (and the missing second row of an odd count) read zero,
        // so all loads are issued together without branches
        const __amdgpu_buffer_rsrc_t ra = __builtin_amdgcn_make_buffer_rsrc(
            const_cast<float *>(J.in + input_row(J, r0) * J.width), (short)0, J.width * 4,
            0x00020000);
        const __amdgpu_buffer_rsrc_t rb = __builtin_amdgcn_make_buffer_rsrc(
            const_cast<float *>(J.in + input_row(J, r1c) * J.width), (short)0,
            has1 ? J.width * 4 : 0, 0x00020000);
        float2 v[M];
#pragma unroll
        for (int m = 0; m < M; ++m) {
            const int off = (ln + 64 * m) * 4;
            v[m] = make_float2(__builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(
                                                             ra, off, 0, 0)),
                               __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(
                                                             rb, off, 0, 0)));
        }
        if (J.fdk) {
#pragma unroll
            for (int m = 0; m < M; ++m) {
                if (64 * m < J.width) {  // wave-uniform: skips all-zero groups
                    const int e = ln + 64 * m;
                    v[m].x *= fdk_weight_at(sdd, vv0, cu, J.delta, e);
                    v[m].y *= fdk_weight_at(sdd, vv1, cu, J.delta, e);
                }
            }
        }
        F::forward(v, B, TwM, Tw64, ln);
        const float *gain = g_gain[J.filter] + N;
#pragma unroll
        for (int j = 0; j < J4; ++j)
#pragma unroll
            for (int p1 = 0; p1 < 8; ++p1) {
                const float g = gain[hi * J4 + j + M * (lo + 8 * p1)];
                float2 &x = v[p1 * J4 + j];
                x = make_float2(x.x * g, x.y * g);
                if ((p1 & 3) == 3) __builtin_amdgcn_sched_barrier(0);
            }
        F::inverse(v, B, TwM, Tw64, ln);

        float *oa, *ob;
        if (J.staged == nullptr) {
            oa = J.out + (size_t)r0 * J.width;
            ob = J.out + (size_t)r1 * J.width;
        } else {
            const size_t vstride = (size_t)J.pitch_u * J.pitch_v;
            oa = J.staged + (size_t)view0 * vstride + (size_t)(iv0 + 1) * J.pitch_u + 1;
            ob = J.staged + (size_t)view1 * vstride + (size_t)(iv1 + 1) * J.pitch_u + 1;
        }
        float ma = 0.0f, mb = 0.0f;
#pragma unroll
        for (int m = 0; m < M; ++m) {
            const int e = ln + 64 * m;
            if (e < J.width) {
                const float x = v[m].x * inv_n, y = v[m].y * inv_n;
                oa[e] = x;
                if (has1) ob[e] = y;
                ma = fmaxf(ma, fabsf(x));
                mb = fmaxf(mb, fabsf(y));
            }
        }
        if (J.row_max) {
            for (int off = 32; off > 0; off >>= 1) {
                ma = fmaxf(ma, __shfl_xor(ma, off));
                mb = fmaxf(mb, __shfl_xor(mb, off));
            }
            if (lane == 0) {
                J.row_max[r0] = ma;
                if (has1) J.row_max[r1] = mb;
            }
        }
    }
}

// Workgroups of the one-wave filter: as many as are resident at once on
// the device (persistent), or one per four row pairs for small jobs.
template <int LOG2N>
static void launch_filter_wave(hipStream_t s, const FilterJob &J, int pairs)
{
    static int resident = 0;  // per process; the device is the same MI355X model throughout
    if (resident == 0) {
        int dev = 0, cus = 0, per_cu = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) ==
                hipSuccess &&
            hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &per_cu, filter_wave_kernel<LOG2N>, 64 * kWaveFftWaves, 0) == hipSuccess &&
            cus > 0 && per_cu > 0)
            resident = cus * per_cu;
        else
            resident = 1024;
    }
    const int want = (pairs + kWaveFftWaves - 1) / kWaveFftWaves;
    filter_wave_kernel<LOG2N><<<want < resident ? want : resident, 64 * kWaveFftWaves, 0, s>>>(J);
}

static int launch_filter_pairs(hipStream_t s, const FilterJob &J, int log2n)
{
    int blocks = (J.rows + 1) / 2;
    if (blocks == 0) return CTR_OK;
    if (CTR_WAVE_FILTER && log2n >= 9 && log2n <= 11) {
        if (log2n == 9) launch_filter_wave<9>(s, J, blocks);
        if (log2n == 10) launch_filter_wave<10>(s, J, blocks);
        if (log2n == 11) launch_filter_wave<11>(s, J, blocks);
        CTR_CHECK_LAUNCH();
        return CTR_OK;
    }
#define CTR_FILT_CASE(L)                                                          \
    case L:                                                                       \
        filter_pairs_kernel<L><<<blocks, kFftThreads, 0, s>>>(J);                 \
        break;
#define CTR_STOCK_CASE(L)                                                         \
    case L:                                                                       \
        filter_stockham_kernel<L><<<blocks, (1 << L) / 8, 0, s>>>(J);             \
        break;
    switch (log2n) {
        CTR_FILT_CASE(0) CTR_FILT_CASE(1) CTR_FILT_CASE(2) CTR_FILT_CASE(3)
        CTR_FILT_CASE(4) CTR_FILT_CASE(5) CTR_FILT_CASE(6) CTR_FILT_CASE(7)
        CTR_FILT_CASE(8)
        CTR_STOCK_CASE(9) CTR_STOCK_CASE(10) CTR_STOCK_CASE(11) CTR_STOCK_CASE(12)
        CTR_STOCK_CASE(13)
    default:
        return fail(CTR_ERR_NOT_POWER_OF_TWO, "FFT length 2^%d unsupported", log2n);
    }
#undef CTR_FILT_CASE
#undef CTR_STOCK_CASE
    CTR_CHECK_LAUNCH();
    return CTR_OK;
}

// ---------------------------------------------------------------------------
// Staged layout helpers
// ---------------------------------------------------------------------------
// Zeroes the border texels of an fp32 staged buffer: rows 0, nv+1, nv+2 of
// every view entirely, columns 0 and nu+1 .. pitch_u-1 of the others.  One
// 64-thread workgroup per staged row (pitch_u - nu <= 35 border columns).
// Zeroes the border of an fp32 staged buffer: 8 staged rows per 256-thread
// workgroup, 32 threads per row.  A border row is zeroed whole; a detector
// row gets column 0 (thread 0) and columns nu + 1 .. pitch_u - 1 (threads
// 1..31, 31 columns apart).
constexpr int kBorderRowsPerBlock = 8;
__global__ __launch_bounds__(256) void zero_border_kernel(float *st, long long rows, int nu,
                                                          int nv, int pitch_u, int pitch_v)
{
    const long long row = (long long)blockIdx.x * kBorderRowsPerBlock + (threadIdx.x >> 5);
    if (row >= rows) return;
    const int j = (int)(threadIdx.x & 31);
    const int r = (int)(row % pitch_v);
    float *p = st + (size_t)row * pitch_u;
    if (r == 0 || r > nv) {
        for (int c = j; c < pitch_u; c += 32) p[c] = 0.0f;
    } else if (j == 0) {
        p[0] = 0.0f;
    } else {
        for (int c = nu + j; c < pitch_u; c += 31) p[c] = 0.0f;
    }
}

// Copies plain projections (either layout) into an fp32 staged buffer,
// borders included.
__global__ void stage_plain_kernel(const float *proj, int layout, int nu, int nv,
                                   int n_views, float *st, int pitch_u, int pitch_v)
{
    long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long total = (long long)pitch_u * pitch_v * n_views;
    if (idx >= total) return;
    int c = (int)(idx % pitch_u);
    int r = (int)((idx / pitch_u) % pitch_v);
    int n = (int)(idx / ((long long)pitch_u * pitch_v));
    float v = 0.0f;
    int iu = c - 1, iv = r - 1;
    if (iu >= 0 && iu < nu && iv >= 0 && iv < nv) {
        size_t src = layout == CTR_LAYOUT_PROJECTIONS
                         ? ((size_t)n * nv + iv) * nu + iu
                         : ((size_t)iv * n_views + n) * nu + iu;
        v = proj[src];
    }
    st[idx] = v;
}

// Per-view max |value| of an fp32 staged buffer: one block per view.
__global__ __launch_bounds__(256) void view_max_kernel(const float *st, size_t view_texels,
                                                       float *view_scale)
{
    const float *p = st + (size_t)blockIdx.x * view_texels;
    float m = 0.0f;
    for (size_t t = threadIdx.x; t < view_texels; t += blockDim.x) m = fmaxf(m, fabsf(p[t]));
    for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
    __shared__ float part[4];
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        m = fmaxf(fmaxf(part[0], part[1]), fmaxf(part[2], part[3]));
        view_scale[blockIdx.x] = m > 0.0f ? m : 1.0f;
    }
}

__device__ inline unsigned short f32_to_bf16_rne(float f)
{
    unsigned u = __float_as_uint(f);
    unsigned lsb = (u >> 16) & 1u;
    u += 0x7fffu + lsb;
    return (unsigned short)(u >> 16);
}

__device__ inline unsigned short f32_to_f16_bits(float f)
{
    _Float16 h = (_Float16)f;
    return __builtin_bit_cast(unsigned short, h);
}

// The upper 16 bits h of a texel whose lower 16 bits are `lo`, chosen so
// that the whole 32-bit word read as an FP32 value is nearest to t: the
// backprojector reads the step of a bf16 value-and-step texel that way,
// without clearing the value's bits below it.  For a fixed `lo` the words
// (h << 16) | lo are ordered like their values, so the nearest is h0 - 1, h0
// or h0 + 1 with h0 the upper half of t; the error is at most half a bf16
// step of t (NaN candidates never compare smaller).
__device__ inline unsigned bf16_step_word(float t, unsigned lo)
{
    const unsigned h0 = __float_as_uint(t) >> 16;
    unsigned best = (h0 << 16) | lo;
    float err = fabsf(__uint_as_float(best) - t);
#pragma unroll
    for (int d = -1; d <= 1; d += 2) {
        const unsigned c = (((h0 + (unsigned)d) & 0xffffu) << 16) | lo;
        const float e = fabsf(__uint_as_float(c) - t);
        if (e < err) {
            best = c;
            err = e;
        }
    }
    return best;
}

// fp32 staged -> 16-bit value-and-step texels: texel (c, r) of a view holds
// a value v = q[r][c] / view_scale rounded to 16 bits in its lower half and
// the step d = q[r + 1][c] / view_scale - v in its upper half (q = 0 past
// the view's last staged row), so one 32-bit load gives a column's two
// vertical taps, v and v + d, and the error of the lower tap is the rounding
// of the step alone.  fp16: d rounded to fp16.  bf16: the upper half is
// chosen so that the whole word, read as FP32, is nearest to d
// (bf16_step_word).  Row 0 (the zero border) holds (0, q[1][c]): the first
// data row is its step.
__device__ inline unsigned pack_vstep(float v, float w, int bf16)
{
    if (bf16) {
        const unsigned hv = f32_to_bf16_rne(v);
        return bf16_step_word(w - __uint_as_float(hv << 16), hv);
    }
    const _Float16 hv = (_Float16)v;
    return (unsigned)__builtin_bit_cast(unsigned short, hv) |
           ((unsigned)f32_to_f16_bits(w - (float)hv) << 16);
}

constexpr int kQuantizeBand = 16;  // staged rows per workgroup of quantize_vstep_kernel

// One workgroup per band of kQuantizeBand staged rows of one view (blockIdx.x
// = view * bands + band), four texels per thread: every row is read once
// into registers, with the row below the band (rows past the view read
// zero).  Only the detector texels (rows 1..nv, columns 1..nu)
// of `st` are read; the border is taken as zero, whatever `st` holds there.
__global__ __launch_bounds__(256) void quantize_vstep_kernel(const float *st, int nu, int nv,
                                                             int pitch_u, int pitch_v,
                                                             float *view_scale, unsigned *out,
                                                             int bf16)
{
    const int bands = (pitch_v + kQuantizeBand - 1) / kQuantizeBand;
    const int n = blockIdx.x / bands;
    const int r0 = (blockIdx.x % bands) * kQuantizeBand;
    const int r1 = min(r0 + kQuantizeBand, pitch_v);
    float vs = view_scale[n];
    if (!(vs > 0.0f)) {
        vs = 1.0f;
        if (threadIdx.x == 0 && r0 == 0) view_scale[n] = 1.0f;
    }
    const float inv = 1.0f / vs;
    const float *src = st + (size_t)n * pitch_v * pitch_u;
    unsigned *dst = out + (size_t)n * pitch_v * pitch_u;
    for (int c = 4 * threadIdx.x; c < pitch_u; c += 4 * blockDim.x) {
        // texel c + i is a detector column when 1 <= c + i <= nu (selected,
        // not multiplied by a 0/1 mask, so that the border is +0 whatever
        // its sign or value in `st`)
        const bool m0 = c >= 1 && c <= nu, m1 = c + 1 <= nu, m2 = c + 2 <= nu, m3 = c + 3 <= nu;
        const bool any = c <= nu;
        auto row = [&](int r) {
            if (!any || r < 1 || r > nv) return make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            const float4 v = *reinterpret_cast<const float4 *>(src + (size_t)r * pitch_u + c);
            return make_float4(m0 ? v.x : 0.0f, m1 ? v.y : 0.0f, m2 ? v.z : 0.0f,
                               m3 ? v.w : 0.0f);
        };
        // all the band's rows (and the one below it) are loaded before any
        // is used, so that their loads are in flight together
        float4 q[kQuantizeBand + 1];
#pragma unroll
        for (int i = 0; i <= kQuantizeBand; ++i) q[i] = row(r0 + i);
#pragma unroll
        for (int i = 0; i < kQuantizeBand; ++i) {
            if (r0 + i >= r1) break;
            const float4 a = q[i], b = q[i + 1];
            uint4 o;
            o.x = pack_vstep(a.x * inv, b.x * inv, bf16);
            o.y = pack_vstep(a.y * inv, b.y * inv, bf16);
            o.z = pack_vstep(a.z * inv, b.z * inv, bf16);
            o.w = pack_vstep(a.w * inv, b.w * inv, bf16);
            *reinterpret_cast<uint4 *>(dst + (size_t)(r0 + i) * pitch_u + c) = o;
        }
    }
}

static inline unsigned grid1d(long long total, int block)
{
    return (unsigned)((total + block - 1) / block);
}

// fp32 staged -> caller's staged buffer in `dtype`.  With `have_max` the
// per-view maxima are already in view_scale (from the filter, which writes the
// detector texels of `st32` only); without, `st32` is whole, border zero
// included (stage_plain_kernel), and view_max_kernel reads all of it.
static int finish_staged(hipStream_t s, const float *st32, int nu, int nv, int n_views,
                         void *staged, int dtype, float *view_scale, bool have_max)
{
    StagedLayout L = staged_layout(nu, nv, n_views, dtype);
    if (dtype == CTR_F32) return CTR_OK;
    CTR_REQUIRE(view_scale, CTR_ERR_NULL_POINTER,
                "view_scale is required for 16-bit staged projections");
    if (!have_max) {
        view_max_kernel<<<n_views, 256, 0, s>>>(st32, L.view, view_scale);
        CTR_CHECK_LAUNCH();
    }
    const long long blocks =
        (long long)((L.pitch_v + kQuantizeBand - 1) / kQuantizeBand) * n_views;
    CTR_REQUIRE(blocks < (1LL << 31), CTR_ERR_INVALID_ARGUMENT, "too many staged rows");
    quantize_vstep_kernel<<<(unsigned)blocks, 256, 0, s>>>(st32, nu, nv, L.pitch_u, L.pitch_v,
                                                           view_scale, (unsigned *)staged,
                                                           dtype == CTR_BF16);
    CTR_CHECK_LAUNCH();
    return CTR_OK;
}

// ---------------------------------------------------------------------------
// Elementwise complex helpers
// ---------------------------------------------------------------------------
__global__ void shift_rows_kernel(const float2 *src, float2 *dst, int n, int rows, int shift)
{
    // one block row-chunk; out-of-place semantics via LDS-free gather:
    // dst[i] = src[(i - shift) mod n]
    long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long total = (long long)n * rows;
    if (idx >= total) return;
    int i = (int)(idx % n);
    long long row = idx / n;
    int sh = shift % n;
    if (sh < 0) sh += n;
    int j = i - sh;
    if (j < 0) j += n;
    dst[idx] = src[row * n + j];
}

// In-place shift of rows of n <= MAXN complex values: one workgroup per row
// holds the row in LDS, so the shift needs no second buffer.
template <int MAXN>
__global__ __launch_bounds__(256) void shift_rows_lds_kernel(float2 *data, int n, int shift)
{
    __shared__ float2 L[MAXN];
    float2 *row = data + (size_t)blockIdx.x * n;
    for (int i = threadIdx.x; i < n; i += blockDim.x) L[i] = row[i];
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        int j = i + shift;  // 0 <= shift < n
        if (j >= n) j -= n;
        row[j] = L[i];
    }
}

// Reverses [a, a + len) of every row in place, one thread per swapped pair:
// three reversals rotate a row longer than the LDS variant holds.
__global__ void reverse_segment_kernel(float2 *data, int n, long long rows, int a, int len)
{
    const long long half = len / 2;
    long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= half * rows) return;
    const long long r = idx / half;
    const int p = (int)(idx - r * half);
    float2 *row = data + (size_t)r * n + a;
    const float2 t = row[p];
    row[p] = row[len - 1 - p];
    row[len - 1 - p] = t;
}

__global__ void blit_kernel(int mode, const float *src, int sw, float *dst, int dw,
                            int w, int h)
{
    long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)w * h) return;
    int x = (int)(idx % w), y = (int)(idx / w);
    size_t si = (size_t)y * sw + x, di = (size_t)y * dw + x;
    switch (mode) {
    case CTR_BLIT_REAL_TO_COMPLEX:
        dst[2 * di] = src[si];
        dst[2 * di + 1] = 0.0f;
        break;
    case CTR_BLIT_COMPLEX_TO_COMPLEX:
        dst[2 * di] = src[2 * si];
        dst[2 * di + 1] = src[2 * si + 1];
        break;
    case CTR_BLIT_COMPLEX_TO_REAL_R:
        dst[di] = src[2 * si];
        break;
    default:
        dst[di] = src[2 * si + 1];
        break;
    }
}

__global__ void conj_scale_kernel(float2 *data, long long count, float conj_sign, float mul)
{
    long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= count) return;
    float2 v = data[idx];
    data[idx] = make_float2(v.x * mul, v.y * conj_sign * mul);
}

__global__ void bit_reverse_kernel(float2 *data, int n, int log2n, long long total)
{
    long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= total) return;
    int i = (int)(idx % n);
    unsigned j = log2n == 0 ? 0u : (__brev((unsigned)i) >> (32 - log2n));
    if ((int)j > i) {
        size_t row = (size_t)(idx / n) * n;
        float2 t = data[row + i];
        data[row + i] = data[row + j];
        data[row + j] = t;
    }
}

__global__ void single_stage_kernel(const float2 *rd, float2 *wr, int n, long long total,
                                    int size)
{
    long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= total) return;
    int i = (int)(idx % n);
    size_t row = (size_t)(idx / n) * n;
    int half = size / 2, local = i % size;
    int k = (local % half) * (n / size);
    float2 w = g_twiddle[k * (kMaxN / n)];
    if (local >= half) {
        float2 a = rd[row + i - half], t = cmul(rd[row + i], w);
        wr[row + i] = make_float2(a.x - t.x, a.y - t.y);
    } else {
        float2 a = rd[row + i], t = cmul(rd[row + i + half], w);
        wr[row + i] = make_float2(a.x + t.x, a.y + t.y);
    }
}

__global__ void scale_real_kernel(float *data, long long count, float mul)
{
    long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx < count) data[idx] *= mul;
}

}  // namespace ctr

using namespace ctr;

extern "C" {

int ctr_fft_radix2(void *stream, float *data, int n, int rows)
{
    clear_error();
    CTR_REQUIRE(data, CTR_ERR_NULL_POINTER, "ctr_fft_radix2: data is NULL");
    CTR_REQUIRE(is_pow2(n), CTR_ERR_NOT_POWER_OF_TWO,
                "[CTAnalyzer Error] FFFT::fft_radix2() is called but texture's width "
                "(%d) is not a power of 2", n);
    CTR_REQUIRE(n <= kMaxN && rows >= 0, CTR_ERR_INVALID_ARGUMENT,
                "ctr_fft_radix2: n=%d rows=%d outside [1, %d] x [0, inf)", n, rows, kMaxN);
    if (rows == 0) return CTR_OK;
    int rc = ensure_twiddles();
    if (rc) return rc;
    return launch_fft_rows<false>(as_stream(stream), (float2 *)data, log2_exact(n), rows);
}

int ctr_inverse_fft_radix2(void *stream, float *data, int n, int rows)
{
    clear_error();
    CTR_REQUIRE(data, CTR_ERR_NULL_POINTER, "ctr_inverse_fft_radix2: data is NULL");
    CTR_REQUIRE(is_pow2(n), CTR_ERR_NOT_POWER_OF_TWO,
                "[CTAnalyzer Error] FFFT::fft_radix2() is called but texture's width "
                "(%d) is not a power of 2", n);
    CTR_REQUIRE(n <= kMaxN && rows >= 0, CTR_ERR_INVALID_ARGUMENT,
                "ctr_inverse_fft_radix2: n=%d rows=%d out of range", n, rows);
    if (rows == 0) return CTR_OK;
    int rc = ensure_twiddles();
    if (rc) return rc;
    return launch_fft_rows<true>(as_stream(stream), (float2 *)data, log2_exact(n), rows);
}

int ctr_fft_shift(void *stream, const float *src, float *dst, int n, int rows, int shift)
{
    clear_error();
    CTR_REQUIRE(src && dst, CTR_ERR_NULL_POINTER, "ctr_fft_shift: NULL buffer");
    CTR_REQUIRE(n > 0 && rows >= 0, CTR_ERR_INVALID_ARGUMENT, "ctr_fft_shift: bad size");
    long long total = (long long)n * rows;
    if (total == 0) return CTR_OK;
    hipStream_t s = as_stream(stream);
    if (src == dst) {
        // in-place overload (FastFourierTransformGPU.cpp:113-119), with no
        // temporary: a row of up to 8192 values is rotated through LDS, a
        // longer one by three in-place reversals (all, then [0, sh), then
        // [sh, n))
        int sh = shift % n;
        if (sh < 0) sh += n;
        if (sh == 0) return CTR_OK;
        float2 *d = (float2 *)dst;
        CTR_REQUIRE(rows < (1 << 30), CTR_ERR_INVALID_ARGUMENT, "ctr_fft_shift: %d rows", rows);
        if (n <= 1024) {
            shift_rows_lds_kernel<1024><<<rows, 256, 0, s>>>(d, n, sh);
        } else if (n <= kMaxN) {
            shift_rows_lds_kernel<kMaxN><<<rows, 256, 0, s>>>(d, n, sh);
        } else {
            const int seg[3][2] = {{0, n}, {0, sh}, {sh, n - sh}};
            for (const auto &g : seg) {
                const long long pairs = (long long)(g[1] / 2) * rows;
                if (pairs == 0) continue;
                reverse_segment_kernel<<<grid1d(pairs, 256), 256, 0, s>>>(d, n, rows, g[0], g[1]);
                CTR_CHECK_LAUNCH();
            }
            return CTR_OK;
        }
        CTR_CHECK_LAUNCH();
        return CTR_OK;
    }
    shift_rows_kernel<<<grid1d(total, 256), 256, 0, s>>>((const float2 *)src, (float2 *)dst, n,
                                                         rows, shift);
    CTR_CHECK_LAUNCH();
    return CTR_OK;
}

int ctr_blit(void *stream, int mode, const float *src, int src_width, int src_height,
             float *dst, int dst_width, int dst_height)
{
    clear_error();
    CTR_REQUIRE(src && dst, CTR_ERR_NULL_POINTER, "ctr_blit: NULL buffer");
    CTR_REQUIRE(mode >= 0 && mode <= 3, CTR_ERR_INVALID_ARGUMENT, "ctr_blit: bad mode %d", mode);
    CTR_REQUIRE(src_width >= 0 && src_height >= 0 && dst_width >= 0 && dst_height >= 0,
                CTR_ERR_INVALID_ARGUMENT, "ctr_blit: negative extent");
    int w = src_width < dst_width ? src_width : dst_width;
    int h = src_height < dst_height ? src_height : dst_height;
    long long total = (long long)w * h;
    if (total == 0) return CTR_OK;
    blit_kernel<<<grid1d(total, 256), 256, 0, as_stream(stream)>>>(mode, src, src_width, dst,
                                                                   dst_width, w, h);
    CTR_CHECK_LAUNCH();
    return CTR_OK;
}

int ctr_conjugate(void *stream, float *data, size_t count)
{
    clear_error();
    CTR_REQUIRE(data || count == 0, CTR_ERR_NULL_POINTER, "ctr_conjugate: NULL buffer");
    if (count == 0) return CTR_OK;
    conj_scale_kernel<<<grid1d((long long)count, 256), 256, 0, as_stream(stream)>>>(
        (float2 *)data, (long long)count, -1.0f, 1.0f);
    CTR_CHECK_LAUNCH();
    return CTR_OK;
}

int ctr_scale_complex(void *stream, float *data, size_t count, float multiplier)
{
    clear_error();
    CTR_REQUIRE(data || count == 0, CTR_ERR_NULL_POINTER, "ctr_scale_complex: NULL buffer");
    if (count == 0) return CTR_OK;
    conj_scale_kernel<<<grid1d((long long)count, 256), 256, 0, as_stream(stream)>>>(
        (float2 *)data, (long long)count, 1.0f, multiplier);
    CTR_CHECK_LAUNCH();
    return CTR_OK;
}

int ctr_fft_bit_reverse(void *stream, float *data, int n, int rows)
{
    clear_error();
    CTR_REQUIRE(data, CTR_ERR_NULL_POINTER, "ctr_fft_bit_reverse: data is NULL");
    CTR_REQUIRE(is_pow2(n), CTR_ERR_NOT_POWER_OF_TWO, "ctr_fft_bit_reverse: %d is not a power of 2", n);
    CTR_REQUIRE(rows >= 0, CTR_ERR_INVALID_ARGUMENT, "ctr_fft_bit_reverse: rows < 0");
    long long total = (long long)n * rows;
    if (total == 0) return CTR_OK;
    bit_reverse_kernel<<<grid1d(total, 256), 256, 0, as_stream(stream)>>>((float2 *)data, n,
                                                                          log2_exact(n), total);
    CTR_CHECK_LAUNCH();
    return CTR_OK;
}

int ctr_fft_single_stage(void *stream, const float *read, float *write, int n, int rows,
                         int step_index)
{
    clear_error();
    CTR_REQUIRE(read && write && read != write, CTR_ERR_INVALID_ARGUMENT,
                "ctr_fft_single_stage: read and write must be distinct, non-NULL buffers");
    CTR_REQUIRE(is_pow2(n) && n <= kMaxN, CTR_ERR_NOT_POWER_OF_TWO,
                "ctr_fft_single_stage: n=%d must be a power of 2 <= %d", n, kMaxN);
    CTR_REQUIRE(rows >= 0 && step_index >= 0 && (2 << step_index) <= n, CTR_ERR_INVALID_ARGUMENT,
                "ctr_fft_single_stage: step %d out of range for n=%d", step_index, n);
    long long total = (long long)n * rows;
    if (total == 0) return CTR_OK;
    int rc = ensure_twiddles();
    if (rc) return rc;
    single_stage_kernel<<<grid1d(total, 256), 256, 0, as_stream(stream)>>>(
        (const float2 *)read, (float2 *)write, n, total, 2 << step_index);
    CTR_CHECK_LAUNCH();
    return CTR_OK;
}

int ctr_scale_real(void *stream, float *data, size_t count, float multiplier)
{
    clear_error();
    CTR_REQUIRE(data || count == 0, CTR_ERR_NULL_POINTER, "ctr_scale_real: NULL buffer");
    if (count == 0) return CTR_OK;
    scale_real_kernel<<<grid1d((long long)count, 256), 256, 0, as_stream(stream)>>>(
        data, (long long)count, multiplier);
    CTR_CHECK_LAUNCH();
    return CTR_OK;
}

int ctr_apply_filter(void *stream, float *spectrum, int n, int rows, int filter)
{
    clear_error();
    CTR_REQUIRE(spectrum, CTR_ERR_NULL_POINTER, "ctr_apply_filter: NULL buffer");
    CTR_REQUIRE(filter >= CTR_RAM_LAK && filter <= CTR_COSINE, CTR_ERR_INVALID_ARGUMENT,
                "[CTAnalyzer Error] FBP2D::apply_filter() unknown filter type %d", filter);
    CTR_REQUIRE(n > 0 && rows >= 0, CTR_ERR_INVALID_ARGUMENT, "ctr_apply_filter: bad size");
    long long total = (long long)n * rows;
    if (total == 0) return CTR_OK;
    apply_filter_kernel<<<grid1d(total, 256), 256, 0, as_stream(stream)>>>(
        (float2 *)spectrum, n, total, filter);
    CTR_CHECK_LAUNCH();
    return CTR_OK;
}

int ctr_filter_rows(void *stream, const float *in, float *out, int width, int rows,
                    int fft_length, int filter)
{
    clear_error();
    CTR_REQUIRE(in && out, CTR_ERR_NULL_POINTER, "ctr_filter_rows: NULL buffer");
    CTR_REQUIRE(width > 0 && rows >= 0, CTR_ERR_INVALID_ARGUMENT,
                "ctr_filter_rows: width=%d rows=%d", width, rows);
    CTR_REQUIRE(filter >= CTR_RAM_LAK && filter <= CTR_COSINE, CTR_ERR_INVALID_ARGUMENT,
                "[CTAnalyzer Error] FBP2D::apply_filter() unknown filter type %d", filter);
    int n = fft_length == 0 ? ceil_pow2(width) : fft_length;
    CTR_REQUIRE(is_pow2(n), CTR_ERR_NOT_POWER_OF_TWO,
                "ctr_filter_rows: FFT length %d is not a power of 2", n);
    CTR_REQUIRE(n >= width && n <= kMaxN, CTR_ERR_INVALID_ARGUMENT,
                "ctr_filter_rows: FFT length %d must lie in [width=%d, %d]", n, width, kMaxN);
    if (rows == 0) return CTR_OK;
    int rc = ensure_twiddles();
    if (rc) return rc;
    FilterJob J{};
    J.in = in;
    J.width = width;
    J.rows = rows;
    J.filter = filter;
    J.fdk = 0;
    J.layout = CTR_LAYOUT_PROJECTIONS;
    J.nv = 1;
    J.n_views = 1;
    J.out = out;
    J.staged = nullptr;
    return launch_filter_pairs(as_stream(stream), J, log2_exact(n));
}

int ctr_filter_projections(void *stream, ctr_workspace *ws, const float *proj, int layout,
                           int nu, int nv, int n_views, int fft_length, int filter,
                           int fdk_weight, double source_distance, double detector_distance,
                           double detector_width, void *staged, int staged_dtype,
                           float *view_scale)
{
    clear_error();
    CTR_REQUIRE(proj && staged, CTR_ERR_NULL_POINTER, "ctr_filter_projections: NULL buffer");
    CTR_REQUIRE(nu > 0 && nv > 0 && n_views > 0, CTR_ERR_INVALID_ARGUMENT,
                "ctr_filter_projections: sizes (%d, %d, %d)", nu, nv, n_views);
    CTR_REQUIRE(layout == CTR_LAYOUT_PROJECTIONS || layout == CTR_LAYOUT_SINOGRAMS,
                CTR_ERR_INVALID_ARGUMENT, "unknown layout %d", layout);
    CTR_REQUIRE(staged_dtype_valid(staged_dtype), CTR_ERR_INVALID_ARGUMENT,
                "unknown staged dtype %d", staged_dtype);
    CTR_REQUIRE(filter >= CTR_RAM_LAK && filter <= CTR_COSINE, CTR_ERR_INVALID_ARGUMENT,
                "[CTAnalyzer Error] FBP2D::apply_filter() unknown filter type %d", filter);
    if (fdk_weight)
        CTR_REQUIRE(source_distance > 0 && detector_distance >= 0 && detector_width > 0,
                    CTR_ERR_INVALID_ARGUMENT, "ctr_filter_projections: bad FDK geometry");
    int n = fft_length == 0 ? ceil_pow2(nu) : fft_length;
    CTR_REQUIRE(is_pow2(n), CTR_ERR_NOT_POWER_OF_TWO, "FFT length %d is not a power of 2", n);
    CTR_REQUIRE(n >= nu && n <= kMaxN, CTR_ERR_INVALID_ARGUMENT,
                "FFT length %d must lie in [nu=%d, %d]", n, nu, kMaxN);
    ctr_workspace *w = nullptr;
    int rc = workspace_resolve(ws, &w);
    if (rc) return rc;
    if ((rc = ensure_twiddles())) return rc;
    hipStream_t s = as_stream(stream);
    StagedLayout L = staged_layout(nu, nv, n_views, CTR_F32);
    float *st32 = nullptr;
    if (staged_dtype == CTR_F32) {
        st32 = (float *)staged;
    } else {
        void *p = nullptr;
        if ((rc = workspace_get(w, SLOT_STAGE_F32, L.bytes, &p))) return rc;
        st32 = (float *)p;
    }
    const long long rows = (long long)L.pitch_v * n_views;
    CTR_REQUIRE(rows < (1LL << 31), CTR_ERR_INVALID_ARGUMENT, "too many staged rows");
    if (staged_dtype == CTR_F32) {
        // the caller's buffer: its border must read zero (the filter writes
        // the detector texels only); 16-bit quantisation reads no border
        zero_border_kernel<<<grid1d(rows, kBorderRowsPerBlock), 256, 0, s>>>(
            st32, rows, nu, nv, L.pitch_u, L.pitch_v);
        CTR_CHECK_LAUNCH();
    }
    // 16-bit storage: the filter records each row's max |q|, reduced per
    // view into view_scale before quantisation
    const bool fused_max = staged_dtype != CTR_F32;
    float *row_max = nullptr;
    if (fused_max) {
        CTR_REQUIRE(view_scale, CTR_ERR_NULL_POINTER,
                    "view_scale is required for 16-bit staged projections");
        void *p = nullptr;
        if ((rc = workspace_get(w, SLOT_VIEW_MAX, sizeof(float) * (size_t)nv * n_views, &p)))
            return rc;
        row_max = (float *)p;
    }
    FilterJob J{};
    J.in = proj;
    J.width = nu;
    J.rows = nv * n_views;
    J.filter = filter;
    J.fdk = fdk_weight ? 1 : 0;
    J.layout = layout;
    J.nv = nv;
    J.n_views = n_views;
    double sdd = source_distance + detector_distance;
    J.sdd2 = sdd * sdd;
    J.delta = detector_width / nu;
    J.out = nullptr;
    J.staged = st32;
    J.pitch_u = L.pitch_u;
    J.pitch_v = L.pitch_v;
    J.row_max = row_max;
    if ((rc = launch_filter_pairs(s, J, log2_exact(n)))) return rc;
    if (fused_max) {
        view_max_from_rows_kernel<<<n_views, 256, 0, s>>>(row_max, nv, view_scale);
        CTR_CHECK_LAUNCH();
    }
    return finish_staged(s, st32, nu, nv, n_views, staged, staged_dtype, view_scale, fused_max);
}

int ctr_stage_projections(void *stream, ctr_workspace *ws, const float *proj, int layout,
                          int nu, int nv, int n_views, void *staged, int staged_dtype,
                          float *view_scale)
{
    clear_error();
    CTR_REQUIRE(proj && staged, CTR_ERR_NULL_POINTER, "ctr_stage_projections: NULL buffer");
    CTR_REQUIRE(nu > 0 && nv > 0 && n_views > 0, CTR_ERR_INVALID_ARGUMENT,
                "ctr_stage_projections: sizes (%d, %d, %d)", nu, nv, n_views);
    CTR_REQUIRE(layout == CTR_LAYOUT_PROJECTIONS || layout == CTR_LAYOUT_SINOGRAMS,
                CTR_ERR_INVALID_ARGUMENT, "unknown layout %d", layout);
    CTR_REQUIRE(staged_dtype_valid(staged_dtype), CTR_ERR_INVALID_ARGUMENT,
                "unknown staged dtype %d", staged_dtype);
    ctr_workspace *w = nullptr;
    int rc = workspace_resolve(ws, &w);
    if (rc) return rc;
    hipStream_t s = as_stream(stream);
    StagedLayout L = staged_layout(nu, nv, n_views, CTR_F32);
    float *st32 = nullptr;
    if (staged_dtype == CTR_F32) {
        st32 = (float *)staged;
    } else {
        void *p = nullptr;
        if ((rc = workspace_get(w, SLOT_STAGE_F32, L.bytes, &p))) return rc;
        st32 = (float *)p;
    }
    long long total = (long long)L.view * n_views;
    stage_plain_kernel<<<grid1d(total, 256), 256, 0, s>>>(proj, layout, nu, nv, n_views, st32,
                                                          L.pitch_u, L.pitch_v);
    CTR_CHECK_LAUNCH();
    return finish_staged(s, st32, nu, nv, n_views, staged, staged_dtype, view_scale, false);
}

}  // extern "C"
