// backproject.hip -- voxel-driven cone / parallel backprojection for gfx950.
//
// One launch covers every view.  A thread owns one (x, z) column and a run of
// consecutive voxels along the rotation axis y, accumulating them in
// registers across all views; the volume is written once at the end.  For a
// fixed column the detector u coordinate, the distance weight and the v step
// are the same for the whole run, so per view a thread sets the column up
// once -- in FP64, from FP64 per-view projection matrices -- and then spends
// 11 FP32 VALU ops per voxel: the v FMA, a clamp, fract + cvt for the row
// and its weight, a 24-bit mad for the texel address, two vertical lerps and
// two pre-weighted accumulations.
//
// Coordinates are "staged coordinates" (one border texel before the data),
// where clamping to [0, n + 1] is the same as "zero outside the detector"
// and truncation equals floor.  The v coordinate of a run is split into an
// integer row I and a fraction f < 1 in FP64, so the per-voxel FP32 value
// f + dv * t stays small and keeps ~1e-5 px resolution even on 1536-row
// panels (a plain FP32 row coordinate has 1.2e-4 px steps there).
//
// A 256-thread block is 4 wave64s, each covering an 8 x 8 patch of columns,
// so the 64 lanes of a wave sample a compact detector footprint.
#include <cstdlib>
#include <cstring>

#include "ctr_internal.h"

namespace ctr {

constexpr int kGeom64 = 12;  // doubles per view: Ui Uk U0 Vj Vi Vk V0 Di Dk D0 - -

// Per-view projective rows (see build_geometry): FP32 copies for the tile
// footprints and FP64 rows for the voxel coordinates.
__global__ void geometry_kernel(float4 *rec, double *rec64, int n_views, int geometry,
                                double rot, double offset, double sod, double sdd, double delta,
                                double d, int X, int Y, int Z, int nu, int nv)
{
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= n_views) return;
    // FBP3D.cpp:536-540: theta_n = 2 pi rot n / N + offset
    double th = 2.0 * M_PI * rot * (double)n / (double)n_views + offset;
    double s, c;
    sincos(th, &s, &c);
    double ox = 0.5 - 0.5 * X, oy = 0.5 - 0.5 * Y, oz = 0.5 - 0.5 * Z;
    double cu = 0.5 * nu - 0.5, cv = 0.5 * nv - 0.5;
    double Di, Dk, D0, Ui, Uk, U0, Vi, Vk, V0, Vj;
    if (geometry == CTR_GEOMETRY_CONE) {
        double mag = sdd / (sod * delta);
        Di = -c * d / sod;
        Dk = -s * d / sod;
        D0 = 1.0 - (c * d * ox + s * d * oz) / sod;
        Ui = -s * d * mag + cu * Di;
        Uk = c * d * mag + cu * Dk;
        U0 = (-s * d * ox + c * d * oz) * mag + cu * D0;
        Vj = d * mag;
        Vi = cv * Di;
        Vk = cv * Dk;
        V0 = d * oy * mag + cv * D0;
    } else {
        Di = 0.0;
        Dk = 0.0;
        D0 = 1.0;
        Ui = -s * d / delta;
        Uk = c * d / delta;
        U0 = (-s * d * ox + c * d * oz) / delta + cu;
        Vj = d / delta;
        Vi = 0.0;
        Vk = 0.0;
        V0 = d * oy / delta + cv;
    }
    rec[3 * n + 0] = make_float4((float)Ui, (float)Uk, (float)U0, (float)Vj);
    rec[3 * n + 1] = make_float4((float)Vi, (float)Vk, (float)V0, 0.0f);
    rec[3 * n + 2] = make_float4((float)Di, (float)Dk, (float)D0, 0.0f);
    double *r = rec64 + (size_t)kGeom64 * n;
    r[0] = Ui;
    r[1] = Uk;
    r[2] = U0;
    r[3] = Vj;
    r[4] = Vi;
    r[5] = Vk;
    r[6] = V0;
    r[7] = Di;
    r[8] = Dk;
    r[9] = D0;
    r[10] = 0.0;
    r[11] = 0.0;
}

int check_geometry(int geometry, double sod, double odd, double detector_width,
                   double volume_width, int X, int Y, int Z, int nu, int nv, int n_views)
{
    CTR_REQUIRE(geometry == CTR_GEOMETRY_CONE || geometry == CTR_GEOMETRY_PARALLEL,
                CTR_ERR_INVALID_ARGUMENT, "unknown geometry %d", geometry);
    CTR_REQUIRE(X > 0 && Y > 0 && Z > 0 && nu > 0 && nv > 0 && n_views > 0,
                CTR_ERR_INVALID_ARGUMENT,
                "sizes must be positive: volume (%d, %d, %d), detector (%d, %d), views %d",
                X, Y, Z, nu, nv, n_views);
    CTR_REQUIRE(X < (1 << 20) && Y < (1 << 20) && Z < (1 << 20) && nu < (1 << 20) &&
                    nv < (1 << 20),
                CTR_ERR_INVALID_ARGUMENT, "sizes must be below 2^20");
    CTR_REQUIRE(detector_width > 0 && volume_width > 0, CTR_ERR_INVALID_ARGUMENT,
                "detector_width and volume_width must be positive");
    if (geometry == CTR_GEOMETRY_CONE) {
        double d = volume_width / X;
        double rx = 0.5 * X * d, rz = 0.5 * Z * d;
        double radius = sqrt(rx * rx + rz * rz);
        CTR_REQUIRE(odd >= 0 && sod > radius * 1.0001, CTR_ERR_INVALID_ARGUMENT,
                    "source distance %.6g must exceed the volume's half diagonal %.6g "
                    "(and detector distance must be >= 0)", sod, radius);
    }
    return CTR_OK;
}

// Detector footprint of every 16 x 16 column tile of the LDS kernel, per
// view, from the tile's 4 corner columns in FP64: (min su, max su, min 1/den,
// max 1/den) in staged coordinates.  For a run of rows j the staged v range
// follows as cv + 1 + Vj (j + oy) [min, max] 1/den, because
// fv = cv + Vj (j + oy) / den for this parameterisation (cone and parallel).
__global__ void footprint_kernel(const double *g64, int n_views, int gx, int gy, float4 *out)
{
    long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long total = (long long)n_views * gx * gy;
    if (idx >= total) return;
    const int bx = (int)(idx % gx);
    const int by = (int)((idx / gx) % gy);
    const int n = (int)(idx / ((long long)gx * gy));
    const double *g = g64 + (size_t)kGeom64 * n;
    double umin = 1e300, umax = -1e300, imin = 1e300, imax = -1e300;
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) {
            const double ci = bx * 16 + 15 * a, ck = by * 16 + 15 * b;
            const double inv = 1.0 / fma(g[7], ci, fma(g[8], ck, g[9]));
            const double su = fma(g[0], ci, fma(g[1], ck, g[2])) * inv + 1.0;
            umin = fmin(umin, su);
            umax = fmax(umax, su);
            imin = fmin(imin, inv);
            imax = fmax(imax, inv);
        }
    out[idx] = make_float4((float)umin, (float)umax, (float)imin, (float)imax);
}

// Bytes of geometry records per view (FP32 footprint rows + FP64 rows).
constexpr size_t kGeomBytesPerView = kGeometryBytesPerView;
static_assert(kGeometryBytesPerView == sizeof(float4) * 3 + sizeof(double) * kGeom64, "layout");

struct GeomPtrs {
    const float4 *f32;
    const double *f64;
};

// Writes the records of `n_views` views at `base` (room for `n_views`
// float4 triples followed by the FP64 rows).
static int build_geometry_at(hipStream_t s, void *base, int geometry, int n_views, double sod,
                             double odd, double detector_width, double volume_width,
                             double rotation_count, double rotation_offset, int X, int Y, int Z,
                             int nu, int nv, GeomPtrs *out)
{
    float4 *f = (float4 *)base;
    double *g = (double *)(f + 3 * (size_t)n_views);
    geometry_kernel<<<(n_views + 63) / 64, 64, 0, s>>>(
        f, g, n_views, geometry, rotation_count, rotation_offset, sod, sod + odd,
        detector_width / nu, volume_width / X, X, Y, Z, nu, nv);
    CTR_CHECK_LAUNCH();
    out->f32 = f;
    out->f64 = g;
    return CTR_OK;
}

int build_geometry(hipStream_t s, ctr_workspace *ws, int geometry, int n_views, double sod,
                   double odd, double detector_width, double volume_width,
                   double rotation_count, double rotation_offset, int X, int Y, int Z, int nu,
                   int nv, const float4 **out)
{
    void *p = nullptr;
    int rc = workspace_get(ws, SLOT_GEOMETRY, kGeomBytesPerView * (size_t)n_views, &p);
    if (rc) return rc;
    GeomPtrs g;
    if ((rc = build_geometry_at(s, p, geometry, n_views, sod, odd, detector_width, volume_width,
                                rotation_count, rotation_offset, X, Y, Z, nu, nv, &g)))
        return rc;
    *out = g.f32;
    return CTR_OK;
}

// Two horizontal taps of one staged row at byte offset `off`, as fp32.
template <int DT>
struct Taps;

template <>
struct Taps<CTR_F32> {
    struct __attribute__((aligned(4))) pair_t {
        float a, b;
    };
    __device__ static inline float2 load(const char *base, unsigned off)
    {
        pair_t p = *reinterpret_cast<const pair_t *>(base + off);
        return make_float2(p.a, p.b);
    }
};

__device__ inline float2 f16_pair(unsigned w)
{
    return make_float2((float)__builtin_bit_cast(_Float16, (unsigned short)(w & 0xffffu)),
                       (float)__builtin_bit_cast(_Float16, (unsigned short)(w >> 16)));
}

__device__ inline float2 bf16_pair(unsigned w)
{
    return make_float2(__uint_as_float(w << 16), __uint_as_float(w & 0xffff0000u));
}

template <>
struct Taps<CTR_F16> {
    __device__ static inline float2 load(const char *base, unsigned off)
    {
        return f16_pair(*reinterpret_cast<const unsigned *>(base + off));
    }
};

template <>
struct Taps<CTR_BF16> {
    __device__ static inline float2 load(const char *base, unsigned off)
    {
        return bf16_pair(*reinterpret_cast<const unsigned *>(base + off));
    }
};

// Column quantities of one view, FP64 from the FP64 projection rows.
struct ColumnView {
    double su;   // staged u coordinate
    double sv0;  // staged v coordinate of the run's first voxel
    float dv;    // v step per voxel along y
    float w2;    // distance weight squared times the scale
};

__device__ inline ColumnView column_view64(const double *g, double fi, double fk, double fj0,
                                           float vs)
{
    ColumnView c;
    const double den = fma(g[7], fi, fma(g[8], fk, g[9]));
    // reciprocal with one Newton step: full FP64 precision
    double r = __builtin_amdgcn_rcp(den);
    r = fma(r, fma(-den, r, 1.0), r);
    c.su = fma(g[0], fi, fma(g[1], fk, g[2])) * r + 1.0;
    c.sv0 = fma(g[3], fj0, fma(g[4], fi, fma(g[5], fk, g[6]))) * r + 1.0;
    c.dv = (float)(g[3] * r);
    c.w2 = (float)(r * r) * vs;
    return c;
}

// Per-voxel constants of one column run: taps at byte offset
// mad24(row, pitch, base) of the staged view (or LDS footprint), row =
// (unsigned)x with x = clamp(f + dv * t, lo, hi) >= 0.
struct Run {
    float cL, cR;   // (1 - wu) w2, wu w2
    float f, dv;    // fractional row of the first voxel, row step
    float lo, hi;   // clamp of x (rows relative to I)
    unsigned base;  // byte offset of row I, column iu
};

// su_range / rows: clamp interval of su and of the absolute row; org_u /
// org_v: texel coordinates of byte offset 0; pitch in bytes.
__device__ inline Run make_run(const ColumnView &c, double su_hi, double org_u, int row_lo,
                               int row_hi, int org_v, unsigned pitch, double su_global_hi)
{
    Run R;
    const double sug = fmin(fmax(c.su, 0.0), su_global_hi);
    const double su = fmin(fmax(sug - org_u, 0.0), su_hi);
    const double sfl = floor(su);
    const float wu = (float)(su - sfl);
    R.cR = c.w2 * wu;
    R.cL = c.w2 - R.cR;
    // rows: keep the run start inside [row_lo - 2^20, row_hi]
    const double sv0 = fmin(fmax(c.sv0, (double)(row_lo - (1 << 20))), (double)row_hi);
    const double ifl = floor(sv0);
    const int I = (int)ifl;
    R.f = (float)(sv0 - ifl);
    R.dv = c.dv;
    R.lo = (float)(row_lo - I);
    R.hi = (float)(row_hi - I);
    R.base = (unsigned)sfl * 4u + (unsigned)(I - org_v) * pitch;
    return R;
}

template <int DT, int VY>
__device__ inline void gather_run(float (&acc)[VY], const Run &R, const char *top,
                                  const char *bot, unsigned pitch)
{
#pragma unroll
    for (int t = 0; t < VY; ++t) {
        const float x = __builtin_amdgcn_fmed3f(fmaf(R.dv, (float)t, R.f), R.lo, R.hi);
        const float wv = __builtin_amdgcn_fractf(x);
        const unsigned off = __umul24((unsigned)x, pitch) + R.base;
        const float2 a = Taps<DT>::load(top, off);
        const float2 b = Taps<DT>::load(bot, off);
        const float left = fmaf(wv, b.x - a.x, a.x);
        const float right = fmaf(wv, b.y - a.y, a.y);
        acc[t] = fmaf(R.cR, right, fmaf(R.cL, left, acc[t]));
    }
}

template <int VY>
__device__ inline void store_run(const float (&acc)[VY], float *vol, int i, int j0, int k, int X,
                                 int Y, int Z, int accumulate, int clip)
{
    if (i >= X || k >= Z) return;
#pragma unroll
    for (int t = 0; t < VY; ++t) {
        const int j = j0 + t;
        if (j < Y) {
            const size_t idx = ((size_t)k * Y + j) * X + i;
            float v = acc[t];
            if (accumulate) v += vol[idx];
            if (clip) v = fmaxf(v, 0.0f);
            vol[idx] = v;
        }
    }
}

// One reconstruction's backprojection arguments; a packed launch carries
// several of them as a kernel argument (no device-side table to upload).
struct BpJob {
    const void *staged;
    const float4 *geom;
    const double *geom64;
    const float4 *foot;  // footprint table (LDS kernel), gx * gy per view
    const float *view_scale;
    float *vol;
    size_t view_texels;
    int pitch_u, pitch_v, nu, nv, n_views;
    int X, Y, Z;
    float scale;
    int accumulate, clip;
    int gx, gy;       // tiles along x and z
    int block_begin;  // first flat block of this job in a packed launch
    // LDS row pitch in bytes, passed at run time: with a compile-time
    // constant the compiler widens the 24-bit row mad into a 64-bit one
    unsigned lds_pitch4;
};

constexpr int kMaxPacked = 16;
struct BpBatch {
    BpJob job[kMaxPacked];
    int count;
};

// ---------------------------------------------------------------------------
// Direct-tap kernel: taps straight from the staged buffer (global loads
// through L1/L2).
// ---------------------------------------------------------------------------
template <int DT, int VY>
__global__ __launch_bounds__(256) void backproject_kernel(BpJob J)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int i = blockIdx.x * 16 + (wave & 1) * 8 + (lane & 7);
    const int k = blockIdx.y * 16 + (wave >> 1) * 8 + (lane >> 3);
    const int j0 = blockIdx.z * VY;

    float acc[VY];
#pragma unroll
    for (int t = 0; t < VY; ++t) acc[t] = 0.0f;

    const char *vbase = reinterpret_cast<const char *>(J.staged);
    const size_t view_bytes = J.view_texels * 4;
    const unsigned pitch4 = (unsigned)J.pitch_u * 4u;
    const double su_hi = (double)J.nu + 1.0;

#pragma unroll 1
    for (int n = 0; n < J.n_views; ++n) {
        float vs = J.scale;
        if (J.view_scale) vs *= J.view_scale[n];
        const ColumnView c =
            column_view64(J.geom64 + (size_t)kGeom64 * n, i, k, j0, vs);
        const Run R = make_run(c, su_hi, 0.0, 0, J.nv + 1, 0, pitch4, su_hi);
        const char *top = vbase + (size_t)n * view_bytes;
        gather_run<DT, VY>(acc, R, top, top + pitch4, pitch4);
    }
    store_run<VY>(acc, J.vol, i, j0, k, J.X, J.Y, J.Z, J.accumulate, J.clip);
}

// ---------------------------------------------------------------------------
// LDS-staged kernel.  Per view, the block's detector footprint -- the
// bounding box of the projections of its 16 x 16 x TY voxel tile, at most
// 63 texels wide and HMAX - 1 rows high -- is copied from the staged
// projections into LDS with global_load_lds (one wave instruction per row,
// no VGPR round trip), and every tap is then an LDS read: two ds_read2_b32
// for fp32, one ds_read2_b32 of two 16-bit pairs otherwise.  A view whose
// footprint does not fit falls back to the direct global taps.
// ---------------------------------------------------------------------------
constexpr int kLdsPitch = 65;  // dwords per LDS row: 64 staged + 1 against bank aliasing
constexpr int kLdsCols = 64;   // texels staged per row (one per lane)

// Taps at LDS byte address `addr`.
template <int DT>
__device__ inline void lds_taps(const unsigned *L, unsigned addr, float2 &top, float2 &bot);

template <>
__device__ inline void lds_taps<CTR_F32>(const unsigned *L, unsigned addr, float2 &top,
                                         float2 &bot)
{
    const float *F = reinterpret_cast<const float *>(reinterpret_cast<const char *>(L) + addr);
    top = make_float2(F[0], F[1]);
    bot = make_float2(F[kLdsPitch], F[kLdsPitch + 1]);
}

template <>
__device__ inline void lds_taps<CTR_F16>(const unsigned *L, unsigned addr, float2 &top,
                                         float2 &bot)
{
    const unsigned *P = reinterpret_cast<const unsigned *>(reinterpret_cast<const char *>(L) + addr);
    top = f16_pair(P[0]);
    bot = f16_pair(P[kLdsPitch]);
}

template <>
__device__ inline void lds_taps<CTR_BF16>(const unsigned *L, unsigned addr, float2 &top,
                                          float2 &bot)
{
    const unsigned *P = reinterpret_cast<const unsigned *>(reinterpret_cast<const char *>(L) + addr);
    top = bf16_pair(P[0]);
    bot = bf16_pair(P[kLdsPitch]);
}

template <int DT, int TY, int HMAX>
__device__ inline void backproject_lds_tile(const BpJob &J, int bx, int by, int bz,
                                            unsigned *lds)
{
    const int pitch_u = J.pitch_u, pitch_v = J.pitch_v, nu = J.nu, nv = J.nv;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int ti0 = bx * 16, tk0 = by * 16;
    const int i = ti0 + (wave & 1) * 8 + (lane & 7);
    const int k = tk0 + (wave >> 1) * 8 + (lane >> 3);
    const int j0 = bz * TY;
    const float fnu = (float)nu, fnv = (float)nv;
    const double su_hi = (double)nu + 1.0;
    // rows of the run relative to the volume centre, and the staged v centre
    const float oy = 0.5f - 0.5f * (float)J.Y;
    const float ja = (float)j0 + oy, jb = (float)(j0 + TY - 1) + oy;
    const float cv1 = 0.5f * fnv + 0.5f;
    const float4 *foot = J.foot + (size_t)by * J.gx + bx;
    const size_t foot_stride = (size_t)J.gx * J.gy;

    float acc[TY];
#pragma unroll
    for (int t = 0; t < TY; ++t) acc[t] = 0.0f;

    const char *vbase = reinterpret_cast<const char *>(J.staged);
    const size_t view_bytes = J.view_texels * 4;
    const unsigned pitch4 = (unsigned)pitch_u * 4u;

#pragma unroll 1
    for (int n = 0; n < J.n_views; ++n) {
        float vs = J.scale;
        if (J.view_scale) vs *= J.view_scale[n];

        // footprint of the tile: u range and 1/den range from the table, the
        // v range of this run of rows from them (0.05 px margin below; the
        // voxels use FP64 coordinates)
        const float4 fp = foot[(size_t)n * foot_stride];
        const float vj = J.geom[3 * n].w;
        const float ya = ja * vj, yb = jb * vj;
        float umin = fp.x, umax = fp.y;
        float vmin = cv1 + fminf(fminf(ya * fp.z, ya * fp.w), fminf(yb * fp.z, yb * fp.w));
        float vmax = cv1 + fmaxf(fmaxf(ya * fp.z, ya * fp.w), fmaxf(yb * fp.z, yb * fp.w));
        umin = __builtin_amdgcn_fmed3f(umin, 0.0f, fnu + 1.0f);
        umax = __builtin_amdgcn_fmed3f(umax, 0.0f, fnu + 1.0f);
        vmin = __builtin_amdgcn_fmed3f(vmin, 0.0f, fnv + 1.0f);
        vmax = __builtin_amdgcn_fmed3f(vmax, 0.0f, fnv + 1.0f);
        const int u0 = __builtin_amdgcn_readfirstlane(max((int)floorf(umin - 0.05f), 0));
        const int u1 = __builtin_amdgcn_readfirstlane(min((int)floorf(umax + 0.05f) + 1, nu + 2));
        const int v0 = __builtin_amdgcn_readfirstlane(max((int)floorf(vmin - 0.05f), 0));
        const int v1 = __builtin_amdgcn_readfirstlane(min((int)floorf(vmax + 0.05f) + 1, nv + 2));
        const int W = u1 - u0 + 1, H = v1 - v0 + 1;

        const ColumnView c = column_view64(J.geom64 + (size_t)kGeom64 * n, i, k, j0, vs);
        const char *top = vbase + (size_t)n * view_bytes;

        if (W < kLdsCols && H < HMAX) {
            // stage staged rows v0 .. v1 + 1 (one spare row, clamped into the view)
            const unsigned *src = reinterpret_cast<const unsigned *>(top);
            const int scol = min(u0 + lane, pitch_u - 1);
            for (int r = wave; r <= H; r += 4) {
                const int srow = min(v0 + r, pitch_v - 1);
                __builtin_amdgcn_global_load_lds(
                    (const void __attribute__((address_space(1))) *)(src + (size_t)srow * pitch_u + scol),
                    (void __attribute__((address_space(3))) *)(lds + r * kLdsPitch), 4, 0, 0);
            }
            asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
            __syncthreads();

            // rows clamped to the staged view and to the footprint
            const unsigned lp4 = J.lds_pitch4;
            const Run R = make_run(c, (double)(W - 1), (double)u0, v0, min(nv + 1, v1), v0, lp4,
                                   su_hi);
#pragma unroll
            for (int t = 0; t < TY; ++t) {
                const float x = __builtin_amdgcn_fmed3f(fmaf(R.dv, (float)t, R.f), R.lo, R.hi);
                const float wv = __builtin_amdgcn_fractf(x);
                const unsigned addr = __umul24((unsigned)x, lp4) + R.base;
                float2 a, b;
                lds_taps<DT>(lds, addr, a, b);
                const float left = fmaf(wv, b.x - a.x, a.x);
                const float right = fmaf(wv, b.y - a.y, a.y);
                acc[t] = fmaf(R.cR, right, fmaf(R.cL, left, acc[t]));
            }
            __syncthreads();
        } else {
            const Run R = make_run(c, su_hi, 0.0, 0, nv + 1, 0, pitch4, su_hi);
            gather_run<DT, TY>(acc, R, top, top + pitch4, pitch4);
        }
    }
    store_run<TY>(acc, J.vol, i, j0, k, J.X, J.Y, J.Z, J.accumulate, J.clip);
}

template <int DT, int TY, int HMAX>
__global__ __launch_bounds__(256) void backproject_lds_kernel(BpJob J)
{
    __shared__ unsigned lds[HMAX * kLdsPitch];
    backproject_lds_tile<DT, TY, HMAX>(J, blockIdx.x, blockIdx.y, blockIdx.z, lds);
}

// Packed launch: the flat block index selects the job (the last job whose
// block_begin it reaches) and the tile inside it.
template <int DT, int TY, int HMAX>
__global__ __launch_bounds__(256) void backproject_lds_packed_kernel(BpBatch B)
{
    __shared__ unsigned lds[HMAX * kLdsPitch];
    const int b = blockIdx.x;
    int j = 0;
#pragma unroll 1
    for (int q = 1; q < B.count; ++q)
        if (b >= B.job[q].block_begin) j = q;
    const BpJob &J = B.job[j];
    const int local = b - J.block_begin;
    const int bx = local % J.gx;
    const int by = (local / J.gx) % J.gy;
    const int bz = local / (J.gx * J.gy);
    backproject_lds_tile<DT, TY, HMAX>(J, bx, by, bz, lds);
}

// Kernel choice: the LDS-staged kernel for volumes at least 32 voxels along
// the rotation axis, the direct-tap kernel otherwise.  CTR_BACKPROJECT_KERNEL
// = "gather" or "lds" overrides (diagnostics and A/B timing).
static int variant_override()
{
    const char *e = getenv("CTR_BACKPROJECT_KERNEL");
    if (!e) return 0;
    if (!strcmp(e, "gather")) return 1;
    if (!strcmp(e, "lds")) return 2;
    return 0;
}

constexpr int kTY = 32, kHMAX = 112;  // LDS kernel: y-run and staged rows

static BpJob make_job(const void *staged, const StagedLayout &L, int nu, int nv, GeomPtrs geom,
                      const float *view_scale, int n_views, float *vol, int X, int Y, int Z,
                      float scale, int accumulate, int clip)
{
    BpJob J;
    J.staged = staged;
    J.geom = geom.f32;
    J.geom64 = geom.f64;
    J.view_scale = view_scale;
    J.vol = vol;
    J.view_texels = L.view;
    J.pitch_u = L.pitch_u;
    J.pitch_v = L.pitch_v;
    J.nu = nu;
    J.nv = nv;
    J.n_views = n_views;
    J.X = X;
    J.Y = Y;
    J.Z = Z;
    J.scale = scale;
    J.accumulate = accumulate;
    J.clip = clip;
    J.gx = (X + 15) / 16;
    J.gy = (Z + 15) / 16;
    J.foot = nullptr;
    J.block_begin = 0;
    J.lds_pitch4 = kLdsPitch * 4u;
    return J;
}

static bool use_lds_kernel(int Y)
{
    const int ov = variant_override();
    return ov == 2 || (ov == 0 && Y >= 32);
}

// Fills the footprint table of job J at `dst` (gx * gy * n_views float4).
static int build_footprints(hipStream_t s, BpJob &J, float4 *dst)
{
    const long long total = (long long)J.gx * J.gy * J.n_views;
    footprint_kernel<<<(unsigned)((total + 255) / 256), 256, 0, s>>>(J.geom64, J.n_views, J.gx,
                                                                      J.gy, dst);
    CTR_CHECK_LAUNCH();
    J.foot = dst;
    return CTR_OK;
}

static size_t footprint_bytes(int X, int Z, int n_views)
{
    return sizeof(float4) * (size_t)((X + 15) / 16) * ((Z + 15) / 16) * n_views;
}

template <int DT>
static int launch_bp(hipStream_t s, const BpJob &J)
{
    if (use_lds_kernel(J.Y)) {
        dim3 grid(J.gx, J.gy, (J.Y + kTY - 1) / kTY);
        backproject_lds_kernel<DT, kTY, kHMAX><<<grid, 256, 0, s>>>(J);
        CTR_CHECK_LAUNCH();
        return CTR_OK;
    }
    const int vy = J.Y >= 16 ? 16 : J.Y >= 8 ? 8 : 1;
#define CTR_BP_CASE(V)                                                                   \
    case V: {                                                                            \
        dim3 grid(J.gx, J.gy, (J.Y + V - 1) / V);                                        \
        backproject_kernel<DT, V><<<grid, 256, 0, s>>>(J);                               \
        break;                                                                           \
    }
    switch (vy) {
        CTR_BP_CASE(1)
        CTR_BP_CASE(8)
        CTR_BP_CASE(16)
    default:
        return fail(CTR_ERR_INVALID_ARGUMENT, "unsupported y-run %d", vy);
    }
#undef CTR_BP_CASE
    CTR_CHECK_LAUNCH();
    return CTR_OK;
}

}  // namespace ctr

using namespace ctr;

extern "C" {

int ctr_backproject(void *stream, ctr_workspace *ws, int geometry, const void *staged,
                    int staged_dtype, const float *view_scale, int nu, int nv, int n_views,
                    double source_distance, double detector_distance, double detector_width,
                    double volume_width, double rotation_count, double rotation_offset,
                    float *volume, int X, int Y, int Z, float scale, int accumulate,
                    int clip_negative)
{
    clear_error();
    CTR_REQUIRE(staged && volume, CTR_ERR_NULL_POINTER,
                "[CTAnalyzer Error] ctr_backproject: projections or volume is NULL");
    CTR_REQUIRE(staged_dtype == CTR_F32 || staged_dtype == CTR_F16 || staged_dtype == CTR_BF16,
                CTR_ERR_INVALID_ARGUMENT, "unknown staged dtype %d", staged_dtype);
    CTR_REQUIRE(staged_dtype == CTR_F32 || view_scale, CTR_ERR_NULL_POINTER,
                "view_scale is required for 16-bit staged projections");
    int rc = check_geometry(geometry, source_distance, detector_distance, detector_width,
                            volume_width, X, Y, Z, nu, nv, n_views);
    if (rc) return rc;
    StagedLayout L = staged_layout(nu, nv, n_views, staged_dtype);
    CTR_REQUIRE(L.view * 4 < ((size_t)1 << 24), CTR_ERR_INVALID_ARGUMENT,
                "staged view of %zu texels exceeds the 2^22-texel limit of the "
                "backprojector's in-view offsets", L.view);
    ctr_workspace *w = nullptr;
    if ((rc = workspace_resolve(ws, &w))) return rc;
    hipStream_t s = as_stream(stream);
    void *p = nullptr;
    if ((rc = workspace_get(w, SLOT_GEOMETRY, kGeomBytesPerView * (size_t)n_views, &p)))
        return rc;
    GeomPtrs g;
    if ((rc = build_geometry_at(s, p, geometry, n_views, source_distance, detector_distance,
                                detector_width, volume_width, rotation_count, rotation_offset, X,
                                Y, Z, nu, nv, &g)))
        return rc;
    BpJob J = make_job(staged, L, nu, nv, g, view_scale, n_views, volume, X, Y, Z, scale,
                       accumulate, clip_negative);
    if (use_lds_kernel(Y)) {
        void *f = nullptr;
        if ((rc = workspace_get(w, SLOT_FOOTPRINT, footprint_bytes(X, Z, n_views), &f))) return rc;
        if ((rc = build_footprints(s, J, (float4 *)f))) return rc;
    }
    switch (staged_dtype) {
    case CTR_F32: return launch_bp<CTR_F32>(s, J);
    case CTR_F16: return launch_bp<CTR_F16>(s, J);
    default: return launch_bp<CTR_BF16>(s, J);
    }
}

int ctr_backproject_packed(void *stream, ctr_workspace *ws, int staged_dtype,
                           const ctr_backproject_job *jobs, int count)
{
    clear_error();
    CTR_REQUIRE(jobs, CTR_ERR_NULL_POINTER, "ctr_backproject_packed: jobs is NULL");
    CTR_REQUIRE(count >= 1 && count <= kMaxPacked, CTR_ERR_INVALID_ARGUMENT,
                "ctr_backproject_packed: count %d outside [1, %d]", count, kMaxPacked);
    CTR_REQUIRE(staged_dtype == CTR_F32 || staged_dtype == CTR_F16 || staged_dtype == CTR_BF16,
                CTR_ERR_INVALID_ARGUMENT, "unknown staged dtype %d", staged_dtype);
    size_t total_views = 0, total_foot = 0;
    for (int q = 0; q < count; q++) {
        const ctr_backproject_job &j = jobs[q];
        total_foot += footprint_bytes(j.X, j.Z, j.n_views);
        CTR_REQUIRE(j.staged && j.volume, CTR_ERR_NULL_POINTER,
                    "ctr_backproject_packed: job %d has a NULL buffer", q);
        CTR_REQUIRE(staged_dtype == CTR_F32 || j.view_scale, CTR_ERR_NULL_POINTER,
                    "ctr_backproject_packed: job %d needs view_scale", q);
        int rc = check_geometry(j.geometry, j.source_distance, j.detector_distance,
                                j.detector_width, j.volume_width, j.X, j.Y, j.Z, j.nu, j.nv,
                                j.n_views);
        if (rc) return rc;
        StagedLayout L = staged_layout(j.nu, j.nv, j.n_views, staged_dtype);
        CTR_REQUIRE(L.view * 4 < ((size_t)1 << 24), CTR_ERR_INVALID_ARGUMENT,
                    "job %d: staged view too large", q);
        total_views += (size_t)j.n_views;
    }
    ctr_workspace *w = nullptr;
    int rc = workspace_resolve(ws, &w);
    if (rc) return rc;
    void *p = nullptr;
    if ((rc = workspace_get(w, SLOT_GEOMETRY, kGeomBytesPerView * total_views, &p))) return rc;
    void *fbase = nullptr;
    if ((rc = workspace_get(w, SLOT_FOOTPRINT, total_foot, &fbase))) return rc;
    hipStream_t s = as_stream(stream);
    BpBatch B;
    B.count = count;
    char *g = (char *)p;
    char *fp = (char *)fbase;
    long long blocks = 0;
    for (int q = 0; q < count; q++) {
        const ctr_backproject_job &j = jobs[q];
        GeomPtrs gp;
        if ((rc = build_geometry_at(s, g, j.geometry, j.n_views, j.source_distance,
                                    j.detector_distance, j.detector_width, j.volume_width,
                                    j.rotation_count, j.rotation_offset, j.X, j.Y, j.Z, j.nu,
                                    j.nv, &gp)))
            return rc;
        StagedLayout L = staged_layout(j.nu, j.nv, j.n_views, staged_dtype);
        B.job[q] = make_job(j.staged, L, j.nu, j.nv, gp, j.view_scale, j.n_views, j.volume, j.X,
                            j.Y, j.Z, j.scale, j.accumulate, j.clip_negative);
        if ((rc = build_footprints(s, B.job[q], (float4 *)fp))) return rc;
        B.job[q].block_begin = (int)blocks;
        blocks += (long long)B.job[q].gx * B.job[q].gy * ((j.Y + kTY - 1) / kTY);
        g += kGeomBytesPerView * (size_t)j.n_views;
        fp += footprint_bytes(j.X, j.Z, j.n_views);
    }
    CTR_REQUIRE(blocks < (1LL << 31), CTR_ERR_INVALID_ARGUMENT, "packed launch too large");
    switch (staged_dtype) {
    case CTR_F32:
        backproject_lds_packed_kernel<CTR_F32, kTY, kHMAX><<<(unsigned)blocks, 256, 0, s>>>(B);
        break;
    case CTR_F16:
        backproject_lds_packed_kernel<CTR_F16, kTY, kHMAX><<<(unsigned)blocks, 256, 0, s>>>(B);
        break;
    default:
        backproject_lds_packed_kernel<CTR_BF16, kTY, kHMAX><<<(unsigned)blocks, 256, 0, s>>>(B);
        break;
    }
    CTR_CHECK_LAUNCH();
    return CTR_OK;
}

}  // extern "C"
