// backproject.hip -- voxel-driven cone / parallel backprojection for gfx950.
//
// One launch covers every view.  A thread owns one (x, z) column and a run of
// VY consecutive voxels along the rotation axis y, accumulating them in
// registers across all views; the volume is written once at the end.  For a
// fixed column the detector u coordinate, the distance weight and the v step
// are the same for the whole run, so per view a thread spends one reciprocal
// and a handful of FMAs on the column and then, per voxel, one FMA for fv,
// a clamp, two 8-byte tap loads (fp32) or two 4-byte pair loads (fp16/bf16)
// and the bilinear blend on the VALU.  Taps read the zero-bordered staged
// layout, so clamping fu to [-1, nu] and fv to [-1, nv] implements "zero
// outside the detector" without branches.
//
// A 256-thread block is 4 wave64s, each covering an 8 x 8 patch of columns,
// so the 64 lanes of a wave sample a compact detector footprint.
#include <cstdlib>
#include <cstring>

#include "ctr_internal.h"

namespace ctr {

// Per-view projective rows, FP64 -> FP32 (see build_geometry).
__global__ void geometry_kernel(float4 *rec, int n_views, int geometry, double rot,
                                double offset, double sod, double sdd, double delta,
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

int build_geometry(hipStream_t s, ctr_workspace *ws, int geometry, int n_views, double sod,
                   double odd, double detector_width, double volume_width,
                   double rotation_count, double rotation_offset, int X, int Y, int Z, int nu,
                   int nv, const float4 **out)
{
    void *p = nullptr;
    int rc = workspace_get(ws, SLOT_GEOMETRY, sizeof(float4) * 3 * (size_t)n_views, &p);
    if (rc) return rc;
    geometry_kernel<<<(n_views + 63) / 64, 64, 0, s>>>(
        (float4 *)p, n_views, geometry, rotation_count, rotation_offset, sod, sod + odd,
        detector_width / nu, volume_width / X, X, Y, Z, nu, nv);
    CTR_CHECK_LAUNCH();
    *out = (const float4 *)p;
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

template <>
struct Taps<CTR_F16> {
    __device__ static inline float2 load(const char *base, unsigned off)
    {
        unsigned w = *reinterpret_cast<const unsigned *>(base + off);
        _Float16 lo = __builtin_bit_cast(_Float16, (unsigned short)(w & 0xffffu));
        _Float16 hi = __builtin_bit_cast(_Float16, (unsigned short)(w >> 16));
        return make_float2((float)lo, (float)hi);
    }
};

template <>
struct Taps<CTR_BF16> {
    __device__ static inline float2 load(const char *base, unsigned off)
    {
        unsigned w = *reinterpret_cast<const unsigned *>(base + off);
        return make_float2(__uint_as_float(w << 16), __uint_as_float(w & 0xffff0000u));
    }
};

template <int DT, int VY>
__global__ __launch_bounds__(256) void backproject_kernel(
    const void *__restrict__ staged, int pitch_u, size_t view_texels, float fnu, float fnv,
    const float4 *__restrict__ geom, const float *__restrict__ view_scale, int n_views,
    float *__restrict__ vol, int X, int Y, int Z, float scale, int accumulate, int clip)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int i = blockIdx.x * 16 + (wave & 1) * 8 + (lane & 7);
    const int k = blockIdx.y * 16 + (wave >> 1) * 8 + (lane >> 3);
    const int j0 = blockIdx.z * VY;
    const float fi = (float)i, fk = (float)k;

    float acc[VY];
#pragma unroll
    for (int t = 0; t < VY; ++t) acc[t] = 0.0f;

    const char *vbase = reinterpret_cast<const char *>(staged);
    const size_t view_bytes = view_texels * 4;
    const unsigned pitch4 = (unsigned)pitch_u * 4u;
    const float fpitch4 = (float)pitch4;
    const float fcol0 = (float)(pitch4 + 4u);  // +1 row, +1 column of border
    const float fj0 = (float)j0;

#pragma unroll 1
    for (int n = 0; n < n_views; ++n) {
        const float4 ga = geom[3 * n + 0];
        const float4 gb = geom[3 * n + 1];
        const float4 gc = geom[3 * n + 2];
        float vs = scale;
        if (view_scale) vs *= view_scale[n];

        const float den = fmaf(gc.x, fi, fmaf(gc.y, fk, gc.z));
        const float inv = __builtin_amdgcn_rcpf(den);
        float fu = fmaf(ga.x, fi, fmaf(ga.y, fk, ga.z)) * inv;
        const float fvb = fmaf(gb.x, fi, fmaf(gb.y, fk, gb.z)) * inv;
        const float dv = ga.w * inv;
        const float w2 = inv * inv * vs;

        fu = __builtin_amdgcn_fmed3f(fu, -1.0f, fnu);
        const float fu0 = floorf(fu);
        const float wu = fu - fu0;
        // byte offset of staged texel (col, row) = ((fv0 + 1) * pitch + fu0 + 1) * 4;
        // every term is an integer below 2^24, so the float arithmetic is exact
        const float fbase = fmaf(fu0, 4.0f, fcol0);
        const float fv_j0 = fmaf(dv, fj0, fvb);
        const char *vptr = vbase + (size_t)n * view_bytes;

#pragma unroll
        for (int t = 0; t < VY; ++t) {
            const float fv = __builtin_amdgcn_fmed3f(fmaf(dv, (float)t, fv_j0), -1.0f, fnv);
            const float fv0 = floorf(fv);
            const float wv = fv - fv0;
            const unsigned off = (unsigned)fmaf(fv0, fpitch4, fbase);
            const float2 top = Taps<DT>::load(vptr, off);
            const float2 bot = Taps<DT>::load(vptr, off + pitch4);
            const float ht = fmaf(wu, top.y - top.x, top.x);
            const float hb = fmaf(wu, bot.y - bot.x, bot.x);
            acc[t] = fmaf(w2, fmaf(wv, hb - ht, ht), acc[t]);
        }
    }

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

// ---------------------------------------------------------------------------
// LDS-staged variant.  Per view, the block's detector footprint -- the
// bounding box of the projections of its 16 x 16 x TY voxel tile, at most
// 63 texels wide and HMAX - 1 rows high -- is copied from the staged
// projections into LDS with global_load_lds (one wave instruction per row,
// no VGPR round trip), and every tap is then an LDS read: two ds_read2_b32
// for fp32, one ds_read2_b32 of two 16-bit pairs otherwise.  A view whose
// footprint does not fit falls back to the direct global taps above.
// ---------------------------------------------------------------------------
constexpr int kLdsPitch = 65;  // dwords per LDS row: 64 staged + 1 against bank aliasing
constexpr int kLdsCols = 64;   // texels staged per row (one per lane)

// Taps at LDS byte address `addr` (row pitch kLdsPitch dwords).
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
    const unsigned a = P[0], b = P[kLdsPitch];
    top = make_float2((float)__builtin_bit_cast(_Float16, (unsigned short)(a & 0xffffu)),
                      (float)__builtin_bit_cast(_Float16, (unsigned short)(a >> 16)));
    bot = make_float2((float)__builtin_bit_cast(_Float16, (unsigned short)(b & 0xffffu)),
                      (float)__builtin_bit_cast(_Float16, (unsigned short)(b >> 16)));
}

template <>
__device__ inline void lds_taps<CTR_BF16>(const unsigned *L, unsigned addr, float2 &top,
                                          float2 &bot)
{
    const unsigned *P = reinterpret_cast<const unsigned *>(reinterpret_cast<const char *>(L) + addr);
    const unsigned a = P[0], b = P[kLdsPitch];
    top = make_float2(__uint_as_float(a << 16), __uint_as_float(a & 0xffff0000u));
    bot = make_float2(__uint_as_float(b << 16), __uint_as_float(b & 0xffff0000u));
}

template <int DT, int TY, int HMAX>
__global__ __launch_bounds__(256) void backproject_lds_kernel(
    const void *__restrict__ staged, int pitch_u, int pitch_v, size_t view_texels, int nu,
    int nv, const float4 *__restrict__ geom, const float *__restrict__ view_scale, int n_views,
    float *__restrict__ vol, int X, int Y, int Z, float scale, int accumulate, int clip)
{
    __shared__ unsigned lds[HMAX * kLdsPitch];

    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int ti0 = blockIdx.x * 16, tk0 = blockIdx.y * 16;
    const int i = ti0 + (wave & 1) * 8 + (lane & 7);
    const int k = tk0 + (wave >> 1) * 8 + (lane >> 3);
    const int j0 = blockIdx.z * TY;
    const float fi = (float)i, fk = (float)k, fj0 = (float)j0;
    const float fnu = (float)nu, fnv = (float)nv;
    // tile corners (voxel-centre indices) for the footprint
    const float ci[2] = {(float)ti0, (float)(ti0 + 15)};
    const float ck[2] = {(float)tk0, (float)(tk0 + 15)};
    const float cj[2] = {fj0, (float)(j0 + TY - 1)};

    float acc[TY];
#pragma unroll
    for (int t = 0; t < TY; ++t) acc[t] = 0.0f;

    const char *vbase = reinterpret_cast<const char *>(staged);
    const size_t view_bytes = view_texels * 4;
    const unsigned pitch4 = (unsigned)pitch_u * 4u;
    const float fpitch4 = (float)pitch4;
    const float fcol0 = (float)(pitch4 + 4u);

#pragma unroll 1
    for (int n = 0; n < n_views; ++n) {
        const float4 ga = geom[3 * n + 0];
        const float4 gb = geom[3 * n + 1];
        const float4 gc = geom[3 * n + 2];
        float vs = scale;
        if (view_scale) vs *= view_scale[n];

        // footprint of the tile: every lane evaluates the 8 corners with the
        // same arithmetic as the voxels below, so the bounds are uniform
        float umin = 3.0e38f, umax = -3.0e38f, vmin = 3.0e38f, vmax = -3.0e38f;
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                const float cden = fmaf(gc.x, ci[a], fmaf(gc.y, ck[b], gc.z));
                const float cinv = __builtin_amdgcn_rcpf(cden);
                const float cu = fmaf(ga.x, ci[a], fmaf(ga.y, ck[b], ga.z)) * cinv;
                const float cvb = fmaf(gb.x, ci[a], fmaf(gb.y, ck[b], gb.z)) * cinv;
                const float cdv = ga.w * cinv;
                umin = fminf(umin, cu);
                umax = fmaxf(umax, cu);
#pragma unroll
                for (int c = 0; c < 2; ++c) {
                    const float cv = fmaf(cdv, cj[c], cvb);
                    vmin = fminf(vmin, cv);
                    vmax = fmaxf(vmax, cv);
                }
            }
        umin = __builtin_amdgcn_fmed3f(umin, -1.0f, fnu);
        umax = __builtin_amdgcn_fmed3f(umax, -1.0f, fnu);
        vmin = __builtin_amdgcn_fmed3f(vmin, -1.0f, fnv);
        vmax = __builtin_amdgcn_fmed3f(vmax, -1.0f, fnv);
        // detector texel ranges (inclusive) holding every tap, 0.01 px margin
        const int u0 = __builtin_amdgcn_readfirstlane(max((int)floorf(umin - 0.01f), -1));
        const int u1 = __builtin_amdgcn_readfirstlane(min((int)floorf(umax + 0.01f) + 1, nu + 1));
        const int v0 = __builtin_amdgcn_readfirstlane(max((int)floorf(vmin - 0.01f), -1));
        const int v1 = __builtin_amdgcn_readfirstlane(min((int)floorf(vmax + 0.01f) + 1, nv + 1));
        const int W = u1 - u0 + 1, H = v1 - v0 + 1;

        const float den = fmaf(gc.x, fi, fmaf(gc.y, fk, gc.z));
        const float inv = __builtin_amdgcn_rcpf(den);
        float fu = fmaf(ga.x, fi, fmaf(ga.y, fk, ga.z)) * inv;
        const float fvb = fmaf(gb.x, fi, fmaf(gb.y, fk, gb.z)) * inv;
        const float dv = ga.w * inv;
        const float w2 = inv * inv * vs;
        fu = __builtin_amdgcn_fmed3f(fu, -1.0f, fnu);
        const char *vptr = vbase + (size_t)n * view_bytes;

        if (W < kLdsCols && H < HMAX) {
            // stage rows v0 .. v1 + 1 (one spare row, clamped into the view)
            const unsigned *src = reinterpret_cast<const unsigned *>(vptr);
            const int scol = min(u0 + 1 + lane, pitch_u - 1);
            for (int r = wave; r <= H; r += 4) {
                const int srow = min(v0 + 1 + r, pitch_v - 1);
                __builtin_amdgcn_global_load_lds(
                    (const void __attribute__((address_space(1))) *)(src + (size_t)srow * pitch_u + scol),
                    (void __attribute__((address_space(3))) *)(lds + r * kLdsPitch), 4, 0, 0);
            }
            asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
            __syncthreads();

            const float ful = __builtin_amdgcn_fmed3f(fu - (float)u0, 0.0f, (float)(W - 1));
            const float fu0 = floorf(ful);
            const float wu = ful - fu0;
            const unsigned colb = (unsigned)fu0 * 4u;
            const float fvl = fmaf(dv, fj0, fvb) - (float)v0;
            const float vhi = (float)(H - 1);
#pragma unroll
            for (int t = 0; t < TY; ++t) {
                const float fv = __builtin_amdgcn_fmed3f(fmaf(dv, (float)t, fvl), 0.0f, vhi);
                const float fv0 = floorf(fv);
                const float wv = fv - fv0;
                const unsigned addr = __umul24((unsigned)fv0, kLdsPitch * 4u) + colb;
                float2 top, bot;
                lds_taps<DT>(lds, addr, top, bot);
                const float ht = fmaf(wu, top.y - top.x, top.x);
                const float hb = fmaf(wu, bot.y - bot.x, bot.x);
                acc[t] = fmaf(w2, fmaf(wv, hb - ht, ht), acc[t]);
            }
            __syncthreads();
        } else {
            const float fu0 = floorf(fu);
            const float wu = fu - fu0;
            const float fbase = fmaf(fu0, 4.0f, fcol0);
            const float fv_j0 = fmaf(dv, fj0, fvb);
#pragma unroll
            for (int t = 0; t < TY; ++t) {
                const float fv = __builtin_amdgcn_fmed3f(fmaf(dv, (float)t, fv_j0), -1.0f, fnv);
                const float fv0 = floorf(fv);
                const float wv = fv - fv0;
                const unsigned off = (unsigned)fmaf(fv0, fpitch4, fbase);
                const float2 top = Taps<DT>::load(vptr, off);
                const float2 bot = Taps<DT>::load(vptr, off + pitch4);
                const float ht = fmaf(wu, top.y - top.x, top.x);
                const float hb = fmaf(wu, bot.y - bot.x, bot.x);
                acc[t] = fmaf(w2, fmaf(wv, hb - ht, ht), acc[t]);
            }
        }
    }

    if (i >= X || k >= Z) return;
#pragma unroll
    for (int t = 0; t < TY; ++t) {
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

template <int DT>
static int launch_bp(hipStream_t s, int vy, const void *staged, const StagedLayout &L, int nu,
                     int nv, const float4 *geom, const float *view_scale, int n_views,
                     float *vol, int X, int Y, int Z, float scale, int accumulate, int clip)
{
    const int ov = variant_override();
    if ((ov == 2 || (ov == 0 && Y >= 32))) {
        constexpr int TY = 32, HMAX = 112;
        dim3 grid((X + 15) / 16, (Z + 15) / 16, (Y + TY - 1) / TY);
        backproject_lds_kernel<DT, TY, HMAX><<<grid, 256, 0, s>>>(
            staged, L.pitch_u, L.pitch_v, L.view, nu, nv, geom, view_scale, n_views, vol, X, Y,
            Z, scale, accumulate, clip);
        CTR_CHECK_LAUNCH();
        return CTR_OK;
    }
    dim3 block(256);
#define CTR_BP_CASE(V)                                                                   \
    case V: {                                                                            \
        dim3 grid((X + 15) / 16, (Z + 15) / 16, (Y + V - 1) / V);                        \
        backproject_kernel<DT, V><<<grid, block, 0, s>>>(                                \
            staged, L.pitch_u, L.view, (float)nu, (float)nv, geom, view_scale, n_views,  \
            vol, X, Y, Z, scale, accumulate, clip);                                      \
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
    const float4 *geom = nullptr;
    if ((rc = build_geometry(s, w, geometry, n_views, source_distance, detector_distance,
                             detector_width, volume_width, rotation_count, rotation_offset, X,
                             Y, Z, nu, nv, &geom)))
        return rc;
    int vy = Y >= 16 ? 16 : (Y >= 8 ? 8 : 1);
    switch (staged_dtype) {
    case CTR_F32:
        return launch_bp<CTR_F32>(s, vy, staged, L, nu, nv, geom, view_scale, n_views, volume,
                                  X, Y, Z, scale, accumulate, clip_negative);
    case CTR_F16:
        return launch_bp<CTR_F16>(s, vy, staged, L, nu, nv, geom, view_scale, n_views, volume,
                                  X, Y, Z, scale, accumulate, clip_negative);
    default:
        return launch_bp<CTR_BF16>(s, vy, staged, L, nu, nv, geom, view_scale, n_views, volume,
                                   X, Y, Z, scale, accumulate, clip_negative);
    }
}

}  // extern "C"
