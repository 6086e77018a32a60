// backproject.hip -- voxel-driven cone / parallel backprojection for gfx950.
//
// One launch covers every view.  A thread owns one (x, z) column and a run of
// consecutive voxels along the rotation axis y, accumulating them in
// registers across all views; the volume is written once at the end.  For a
// fixed column the detector u coordinate, the distance weight and the v step
// are the same for the whole run, so per view a thread sets the column up
// once -- in FP64, from FP64 per-view projection matrices -- and then spends
// a handful of VALU ops per voxel: the v FMA, fract + cvt for the row and
// its weight, a 24-bit mad for the texel address, the two vertical lerps and
// the pre-weighted accumulation (details with each kernel below).
//
// Coordinates are "staged coordinates" (one border texel before the data),
// where clamping to [0, n + 1] is the same as "zero outside the detector"
// and truncation equals floor.  The v coordinate of a run is taken relative
// to an integer row in FP64, so the per-voxel FP32 value f + dv * t stays
// small and keeps ~1e-5 px resolution even on 1536-row panels (a plain FP32
// row coordinate has 1.2e-4 px steps there).
//
// A 256-thread block is 4 wave64s, each covering an 8 x 8 patch of columns,
// so the 64 lanes of a wave sample a compact detector footprint.
#include <mutex>

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

// Detector footprint of every tile x tile column tile of the LDS kernel, per
// view, from the tile's 4 corner columns in FP64: (min su, max su, min 1/den,
// max 1/den) in staged coordinates.  For a run of rows j the staged v range
// follows as cv + 1 + Vj (j + oy) [min, max] 1/den, because
// fv = cv + Vj (j + oy) / den for this parameterisation (cone and parallel).
__global__ void footprint_kernel(const double *g64, int n_views, int gx, int gy, int tile,
                                 float4 *out)
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
            const double ci = bx * tile + (tile - 1) * a, ck = by * tile + (tile - 1) * b;
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

typedef _Float16 half2_t __attribute__((ext_vector_type(2)));

// Two horizontal fp32 taps of one staged row at byte offset `off`.
__device__ inline float2 f32_taps(const char *base, unsigned off)
{
    struct __attribute__((aligned(4))) pair_t {
        float a, b;
    };
    const pair_t p = *reinterpret_cast<const pair_t *>(base + off);
    return make_float2(p.a, p.b);
}

// v + wv d of a 16-bit value-and-step texel (value v in the lower half, step
// d to the next row in the upper half), in FP32.  fp16: both halves widened
// inside one mixed-precision FMA (v_fma_mix_f32, exact widening).  bf16: the
// word itself, read as FP32, is the step (its upper half was chosen for that
// when the texel was written) and the value is the lower half shifted up.
template <int DT>
__device__ inline float vstep_lerp(unsigned w, float wv)
{
    if constexpr (DT == CTR_BF16) {
        return fmaf(wv, __uint_as_float(w), __uint_as_float(w << 16));
    } else {
        const half2_t h = __builtin_bit_cast(half2_t, w);
        return fmaf(wv, (float)h.y, (float)h.x);
    }
}

// Column quantities of one view, FP64 from the FP64 projection rows.
struct ColumnView {
    double su;   // staged u coordinate
    double sv0;  // staged v coordinate of the run's first voxel
    float dv;    // v step per voxel along y
    float w2;    // distance weight squared times the scale
};

// Per-view tables are read-only for the whole launch: reading them through
// the constant address space lets wave-uniform reads use scalar loads,
// which stay off the vector-memory counter that the LDS row prefetch uses.
template <typename T>
__device__ inline const __attribute__((address_space(4))) T *as_constant(const T *p)
{
    return (const __attribute__((address_space(4))) T *)p;
}

__device__ inline ColumnView column_view64(const double *gp, double fi, double fk, double fj0,
                                           float vs)
{
    const __attribute__((address_space(4))) double *g = as_constant(gp);
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

// Direct taps of one column run: fp32 reads rows I and I + 1 (columns iu,
// iu + 1) of the staged view at `top`; the 16-bit value-and-step texels need
// row I only.
template <int DT, int VY>
__device__ inline void gather_run(float (&acc)[VY], const Run &R, const char *top,
                                  unsigned pitch)
{
    if constexpr (staged_is_16bit(DT)) {
#pragma unroll
        for (int t = 0; t < VY; ++t) {
            const float x = __builtin_amdgcn_fmed3f(fmaf(R.dv, (float)t, R.f), R.lo, R.hi);
            const float wv = __builtin_amdgcn_fractf(x);
            const unsigned off = __umul24((unsigned)x, pitch) + R.base;
            const unsigned *P = reinterpret_cast<const unsigned *>(top + off);
            const float left = vstep_lerp<DT>(P[0], wv);
            const float right = vstep_lerp<DT>(P[1], wv);
            acc[t] = fmaf(R.cR, right, fmaf(R.cL, left, acc[t]));
        }
    } else {
#pragma unroll
        for (int t = 0; t < VY; ++t) {
            const float x = __builtin_amdgcn_fmed3f(fmaf(R.dv, (float)t, R.f), R.lo, R.hi);
            const float wv = __builtin_amdgcn_fractf(x);
            const unsigned off = __umul24((unsigned)x, pitch) + R.base;
            const float2 a = f32_taps(top, off);
            const float2 b = f32_taps(top + pitch, off);
            const float left = fmaf(wv, b.x - a.x, a.x);
            const float right = fmaf(wv, b.y - a.y, a.y);
            acc[t] = fmaf(R.cR, right, fmaf(R.cL, left, acc[t]));
        }
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
// through L1/L2), coordinates clamped to the staged view: per voxel the v
// FMA, a clamp, fract, cvt, mad24, two loads, two lerps and two FMAs (11
// VALU for fp32, 9 for fp16).  Used for short runs (Y < 32, the 2D fan / parallel
// workloads) and as the reference variant of the LDS kernel.
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
        gather_run<DT, VY>(acc, R, top, pitch4);
    }
    store_run<VY>(acc, J.vol, i, j0, k, J.X, J.Y, J.Z, J.accumulate, J.clip);
}

// ---------------------------------------------------------------------------
// LDS-staged kernel.
//
// Per view, the block's detector footprint -- the bounding rectangle of the
// projections of its 16 x 16 x TY voxel tile, at most CMAX texels wide and
// HMAX rows high -- is copied from the staged projections into LDS with
// global_load_lds (one wave instruction per row, addresses in SGPRs, no VGPR
// round trip).  Two row buffers alternate: the rows of view n + 1 are in
// flight while view n is interpolated from the other buffer, and one barrier
// per view is the only synchronisation.
//
// The rectangles are computed kRectChunk views at a time into a table in LDS
// (one thread per view) from the per-column-tile footprint table, and each
// is classified as
//   skip   -- every tap lies in the zero border: no contribution, no work;
//   lds    -- staged and interpolated from LDS;
//   direct -- too large for the buffers: done after the main loop by a
//             second pass over such views with direct global taps.
// Rectangles are not clipped to the detector.  Rows and columns outside it
// are loaded from the (zero) border of the staged view by clamping the
// source row and column, so the per-voxel row coordinate needs no clamp.
//
// Per voxel, all arithmetic FP32:
//   fp32 -- v FMA, fract, cvt, mad24, the taps (two ds_read2_b32: rows I and
//     I + 1, columns u and u + 1), two subtractions and two lerps for the
//     columns, two FMAs into the accumulator with the column's (1 - wu, wu)
//     weights (10 VALU, six of them FMA-class);
//   fp16 -- v FMA, fract, cvt, mad24, one ds_read2_b32 of the value-and-step
//     texels of columns u and u + 1 in row I, each column's vertical lerp
//     v + wv d as one v_fma_mix_f32 straight from the texel's f16 halves,
//     the two FMAs into the accumulator (8 VALU, four FMA-class);
//   bf16 -- as fp16, with one shift per texel to widen the value (the step
//     is the texel read as FP32): 10 VALU, four FMA-class.
// The LDS reads are software-pipelined: those of the next batch of voxels
// are issued before the current batch is combined.
// ---------------------------------------------------------------------------
// LDS geometry per staged format: staged columns, row pitch in dwords and
// rows per buffer.  A 16 x 16 column tile spans at most 44 detector columns
// on the 512^3 scan and fewer than 48 on 99.8% of the 1440^3 scan's
// (tile, view) pairs, so a row holds 48 texels (the rest take the
// direct-tap pass).  The pitch adds padding that staggers the banks of
// neighbouring rows.  The lanes of a wave read taps from an 8 x 8 patch of
// columns, i.e. from texels within about +-10 columns and +-2 rows of each
// other; with 32 banks a row shift of s banks makes (row + dr, col + du)
// collide when s dr + du = 0 mod 32: s = 1 (pitch 65) collides for every
// (1, -1) neighbour, s = 5 (69) for |du| >= 5 pairs two rows apart, s = 11
// (75) and s = -11 (53) only for |du| >= 10 or three rows apart.  Two
// buffers of `rows` rows bound the workgroups per CU (160 KB of LDS): fp32
// at 53 x 84 fits four (its 120 VGPRs allow no more), 16-bit texels at
// 51 x 72 five.  Measured (GUPS, cols/pitch/rows; 512^3 x 720, 1024^3 x 1200,
// 1440^3 x 1440; the 16-bit rows with the earlier (q[u], q[u + 1]) pair
// texels, two rows read per voxel, and 16-bit arithmetic):
//   fp32  64/65/96 1712 (512^3 only), 64/75/84 2825 3399 3480,
//         48/53/84 2909 3637 3795, 48/51/84 2805 3625 3804
//   fp16  64/65/96 2664 (512^3 only), 64/69/70 3618 4149 4009,
//         48/53/70 3738 4296 4259, 48/51/72 3760 4302 4351
// fp32 reads four taps per voxel and is the most sensitive to conflicts;
// 16-bit texels read two and profit more from the fifth workgroup.  Every
// 32-voxel footprint of the 512^3 scan fits 72 rows bar 0.01% (98.9% of
// the 1440^3 scan); larger ones take the direct-tap pass.
#ifndef CTR_LDS_COLS32
#define CTR_LDS_COLS32 48
#endif
#ifndef CTR_LDS_PITCH32
#define CTR_LDS_PITCH32 53
#endif
#ifndef CTR_LDS_ROWS32
#define CTR_LDS_ROWS32 84
#endif
#ifndef CTR_LDS_COLS16
#define CTR_LDS_COLS16 48
#endif
#ifndef CTR_LDS_PITCH16
#define CTR_LDS_PITCH16 51
#endif
// 16-bit rows: 60 (six workgroups per CU) measured FP16 1.1% faster at
// 1024^3 x 1200 but 20% slower at 512^3 x 720 and 13% at 1440^3 x 1440, whose
// taller footprints then take the direct-tap pass more often.
#ifndef CTR_LDS_ROWS16
#define CTR_LDS_ROWS16 72
#endif
// Voxels per lane along y.  24 (16-bit rows 54 or 56: 6 workgroups per CU
// instead of 5) measured fp16 3194-3213 / 3692 / 3373-3564 GUPS at 512^3 /
// 1024^3 / 1440^3 against 3366 / 3853 / 3962 for 32: the per-view set-up,
// shared by fewer voxels, costs more than the sixth workgroup gains.  40
// (16-bit rows 80 or 88: 4 workgroups per CU) measured fp16 3256-3363 /
// 3757-3786 / 3586-3757 against 3391 / 3819 / 3948 for 32.
#ifndef CTR_LDS_TY
#define CTR_LDS_TY 32
#endif
// Voxels per software-pipelined batch of LDS reads.  Measured on 512^3 x
// 720 and 1024^3 x 1200 (GUPS): fp32 batch 4 2767 / 3254, batch 2 2792 /
// 3389, batch 8 2625 / 3071 (later, on the current kernel: batch 1 2937 /
// 3635 against batch 2 2943 / 3641).
#ifndef CTR_F32_BATCH
#define CTR_F32_BATCH 2
#endif
// The same for the 16-bit value-and-step texels: batch 2 against batch 1,
// kernel ms at 512^3 / 1024^3 / 1440^3 (same box): fp16 22.20 / 269.5 /
// 801.2 against 22.32-22.48 / 272.0-272.5 / 805.5-810.1, bf16 23.86 / 286.6
// / 835.9 against 23.93-24.07 / 287.6-289.5 / 840.0-841.0.
#ifndef CTR_16_BATCH
#define CTR_16_BATCH 2
#endif
constexpr int kRectChunk = 128; // views per refill of the rectangle table

// Column tile edge of the LDS kernel per format: a tile of T x T columns is
// (T / 8)^2 waves, each owning an 8 x 8 patch.
#ifndef CTR_LDS_TILE32
#define CTR_LDS_TILE32 16
#endif
#ifndef CTR_LDS_TILE16
#define CTR_LDS_TILE16 16
#endif
template <int DT>
struct LdsConfig {
    static constexpr int kCols = CTR_LDS_COLS16, kPitch = CTR_LDS_PITCH16, kRows = CTR_LDS_ROWS16,
                         kTile = CTR_LDS_TILE16;
};
template <>
struct LdsConfig<CTR_F32> {
    static constexpr int kCols = CTR_LDS_COLS32, kPitch = CTR_LDS_PITCH32, kRows = CTR_LDS_ROWS32,
                         kTile = CTR_LDS_TILE32;
};

enum RectMode { RECT_SKIP = 0, RECT_LDS = 1, RECT_DIRECT = 2 };

// A rectangle is int4 (u0, v0, rows, mode | cols << 8).
__device__ inline int rect_mode(const int4 &r) { return r.w & 0xff; }
__device__ inline int rect_cols(const int4 &r) { return r.w >> 8; }

// LDS of the kernel: the rectangle table (kRectChunk + 1 int4), then two row
// buffers.
template <int DT>
struct LdsLayout {
    static constexpr int kCols = LdsConfig<DT>::kCols, kPitch = LdsConfig<DT>::kPitch,
                         kRows = LdsConfig<DT>::kRows, kTile = LdsConfig<DT>::kTile;
    static constexpr int kWaves = (kTile / 8) * (kTile / 8), kThreads = 64 * kWaves;
    static constexpr int kRectBytes = (kRectChunk + 1) * 16;
    static constexpr int kBufDwords = kRows * kPitch;
    static constexpr int kDwords = kRectBytes / 4 + 2 * kBufDwords;
    static_assert((2 * kBufDwords) % 4 == 0, "buffers are cleared as uint4");
    static_assert(kPitch >= kCols + 1 && kPitch + 1 <= 255, "ds_read2 offsets");
    static_assert(kTile % 8 == 0 && kThreads <= 1024 && kThreads > kRectChunk, "tile");
};

// fp32 taps at LDS byte address `addr`: rows I (top) and I + 1 (bot),
// columns (u, u + 1).
__device__ inline void lds_taps_f32(const unsigned *L, unsigned addr, float2 &top, float2 &bot)
{
    constexpr int P = LdsConfig<CTR_F32>::kPitch;
    const float *F = reinterpret_cast<const float *>(reinterpret_cast<const char *>(L) + addr);
    top = make_float2(F[0], F[1]);
    bot = make_float2(F[P], F[P + 1]);
}

// Detector rectangle of one tile for view n: (u0, v0, rows, mode), staged
// coordinates.  Bounds are clamped to [-1024, n + 1024] first so the int
// conversions stay exact; a clamped rectangle is still classified correctly
// because the clamp margin exceeds both buffer limits.  `below` is 1 when
// the format's taps include row I + 1 (fp32), 0 when a texel of row I
// carries both vertical taps (16-bit value-and-step texels).
__device__ inline int4 tile_rect(const BpJob &J, const float4 *foot, size_t foot_stride, int n,
                                 float ja, float jb, float cv1, int hmax, int cmax, int below)
{
    const float4 fp = foot[(size_t)n * foot_stride];
    const float vj = J.geom[3 * n].w;
    const float ya = ja * vj, yb = jb * vj;
    const float lo = fminf(fminf(ya * fp.z, ya * fp.w), fminf(yb * fp.z, yb * fp.w));
    const float hi = fmaxf(fmaxf(ya * fp.z, ya * fp.w), fmaxf(yb * fp.z, yb * fp.w));
    const float ul = (float)J.nu + 1024.0f, vl = (float)J.nv + 1024.0f;
    const int u0 = (int)floorf(__builtin_amdgcn_fmed3f(fp.x - 0.05f, -1024.0f, ul));
    const int u1 = (int)floorf(__builtin_amdgcn_fmed3f(fp.y + 0.05f, -1024.0f, ul)) + 1;
    const int v0 = (int)floorf(__builtin_amdgcn_fmed3f(cv1 + lo - 0.05f, -1024.0f, vl));
    const int v1 = (int)floorf(__builtin_amdgcn_fmed3f(cv1 + hi + 0.05f, -1024.0f, vl)) + below;
    // skip: every tap in the zero border.  Staged row 0 is a zero row of
    // fp32, but a 16-bit texel there holds (0, first data row), so for the
    // 16-bit formats only rows < 0 are zero above the data.
    int mode;
    if (u1 <= 0 || v1 < below || u0 >= J.nu + 1 || v0 >= J.nv + 1)
        mode = RECT_SKIP;
    else if (u1 - u0 < cmax && v1 - v0 < hmax)
        mode = RECT_LDS;
    else
        mode = RECT_DIRECT;
    // start staged rows on a 32-texel (128-byte) boundary when the rectangle
    // still fits the cmax staged columns, else on 16 texels: each row DMA then
    // touches whole cache lines (the staged pitch is a multiple of 32 texels)
    int ua = u0;
    if (mode == RECT_LDS) {
        if (u1 - (u0 & ~31) < cmax)
            ua = u0 & ~31;
        else if (u1 - (u0 & ~15) < cmax)
            ua = u0 & ~15;
    }
    return make_int4(ua, v0, v1 - v0 + 1, mode | ((u1 - ua + 1) << 8));
}

__device__ inline int4 read_rect(const int4 *tab, int c)
{
    const int4 r = tab[c];
    return make_int4(__builtin_amdgcn_readfirstlane(r.x), __builtin_amdgcn_readfirstlane(r.y),
                     __builtin_amdgcn_readfirstlane(r.z), __builtin_amdgcn_readfirstlane(r.w));
}

// Rows v0 .. v0 + h - 1 of one staged view into an LDS row buffer, row r by
// wave r mod 4, `cols` columns from u0 (one per lane).  Rows outside the
// view are read from its last row, columns outside it from its last
// column: both zero in every texel format (row 0 is not -- a 16-bit
// value-and-step texel there holds (0, first data row)).  The loads go
// through a buffer
// resource spanning exactly the view (raw 32-bit format), with the row
// offset in an SGPR and the lane's column in the VGPR offset: no VALU work
// per row, and any address outside the view would read zero.
template <int PITCH, int CMAX, int NW>
__device__ inline void stage_rect(const char *view, unsigned view_bytes, int pitch_u,
                                  int pitch_v, int u0, int v0, int h, int cols, unsigned *buf,
                                  int wave, int lane)
{
    const __amdgpu_buffer_rsrc_t rsrc =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<char *>(view), (short)0, (int)view_bytes,
                                          0x00020000);
    // one load instruction per row and 64 columns; columns outside the view
    // read its column 0 or its last column, both zero in every format
    constexpr int kChunks = (CMAX + 63) / 64;
    unsigned col4[kChunks];
#pragma unroll
    for (int c = 0; c < kChunks; ++c)
        col4[c] = (unsigned)min(max(u0 + 64 * c + lane, 0), pitch_u - 1) * 4u;
    const unsigned row_bytes = (unsigned)pitch_u * 4u;
    for (int r = wave; r < h; r += NW) {
        const int srow = v0 + r < 0 ? pitch_v - 1 : min(v0 + r, pitch_v - 1);
#pragma unroll
        for (int c = 0; c < kChunks; ++c)
            if (64 * c + lane < cols)  // only the rectangle's columns are staged
                __builtin_amdgcn_raw_ptr_buffer_load_lds(
                    rsrc, (__attribute__((address_space(3))) void *)(buf + r * PITCH + 64 * c),
                    4, col4[c], (unsigned)srow * row_bytes, 0, 0);
    }
}

// Row coordinates of the FP16 LDS runs carry a bias of 128: a row x in
// [0, 128) is held as x + 128, a float in [128, 256), whose bits are
// 0x43000000 + floor(x) * 2^16 + the fraction in 16 bits.  The high half of
// the bits is then 0x4300 + floor(x), and one v_mad_u32_u16 on it (op_sel
// selects the high half) gives the LDS address, where an unbiased row needs
// a conversion and a 24-bit multiply-add; v_fract_f32 of the biased row is
// the row weight as before.  The row is then rounded to 2^-16 (1.5e-5
// rows) in place of 2^-17 .. 2^-18 for rows 32 .. 127.  FP32 and BF16 runs
// stay unbiased: FP32 keeps the finer rounding, and a BF16 step carries 8
// bits, so a sample that the coarser rounding moves across a row boundary
// (where v + d of one row and the next row's v differ by the step's
// rounding) would cost BF16 four times what it costs FP16.  Measured
// (kernel ms, 512^3 / 1024^3, alternating runs on one box): FP16
// 21.34-21.39 / 255.1-256.0 against 22.18-22.24 / 270.0-270.1; BF16 biased
// 23.26-23.27 / 279.4-280.4 against 23.74-24.07 / 287.4, but 4.7e-4 max-rel
// against the oracle of its dequantized texels on white noise (4e-4
// unbiased).
#ifndef CTR_ROW_BIAS
#define CTR_ROW_BIAS 1
#endif
template <int DT>
__device__ constexpr bool row_biased()
{
    return CTR_ROW_BIAS && DT == CTR_F16;
}

// LDS byte address of row floor(x) of a run: `base` is the run's address of
// row 0 as lds_base gives it.
template <int DT>
__device__ inline unsigned lds_row_addr(float x, unsigned lp4, unsigned base)
{
    if constexpr (row_biased<DT>()) {
        unsigned a;
        asm("v_mad_u32_u16 %0, %1, %2, %3 op_sel:[1,0,0,0]"
            : "=v"(a)
            : "v"(__float_as_uint(x)), "s"(lp4), "v"(base));
        return a;
    } else {
        return __umul24((unsigned)x, lp4) + base;
    }
}

// The run's row-0 address as lds_row_addr takes it (a biased row's high
// half counts 0x4300 rows too many; the address wraps modulo 2^32).
template <int DT>
__device__ inline unsigned lds_base(unsigned base, unsigned lp4)
{
    return row_biased<DT>() ? base - 0x4300u * lp4 : base;
}

// Per-view constants of a column run read from LDS.
struct LdsRun {
    float cL, cR;   // (1 - wu) w2, wu w2
    float f, dv;    // row of the first voxel relative to the rectangle (biased for FP16), row step
    unsigned base;  // LDS byte address of (row 0, column of the left tap)
};

template <int DT>
__device__ inline LdsRun lds_run(const ColumnView &c, int u0, int v0, unsigned buf_byte, int cmax)
{
    LdsRun R;
    const double su = c.su - (double)u0;
    const double sfl = floor(su);
    const float wu = (float)(su - sfl);
    R.cR = c.w2 * wu;
    R.cL = c.w2 - R.cR;
    R.f = (float)(c.sv0 - (double)v0 + (row_biased<DT>() ? 128.0 : 0.0));
    R.dv = c.dv;
    R.base = buf_byte + (unsigned)min(max((int)sfl, 0), cmax - 2) * 4u;
    return R;
}

// The same run straight from the view's FP64 rows, with the rectangle's
// origin, the staged border's +1 and the FP16 row bias folded into the
// addend of the FMA that applies the reciprocal (one rounding each for the
// column and the row relative to the rectangle, where column_view64 and
// lds_run take three).  FP16 splits its column (below 48 inside the
// rectangle) in FP32, to 2^-19 of a texel.  Measured (kernel ms, 512^3 /
// 1024^3, three alternating runs each on one box) against column_view64 +
// lds_run: FP16 20.69-20.80 / 250.5-252.0 against 20.95-21.12 / 252.7-254.1,
// BF16 23.05-23.30 / 279.8-281.1 against 23.23-23.38 / 282.3-283.3, FP32
// alike at 512^3 and 0.5% faster at 1024^3.
#ifndef CTR_FOLD_ORIGIN
#define CTR_FOLD_ORIGIN 1
#endif
#ifndef CTR_MOV64
#define CTR_MOV64 1
#endif
// A wave-uniform FP64 value (in SGPRs) moved into a VGPR pair in one
// instruction, to start a per-lane FMA chain: the compiler's own copies of
// SGPR pairs are two 32-bit moves.  With this, the LDS pitch taken as an
// SGPR operand and the two scale factors applied from SGPRs, the set-up
// issues 4 fewer VALU instructions per view and lane: FP16 kernel 20.86-20.95
// / 251.0-251.6 ms against 21.12-21.19 / 252.6-254.0 at 512^3 / 1024^3
// (alternating runs on one box), FP32 alike.
__device__ inline double uniform_to_vgpr64(double s)
{
#if CTR_MOV64
    double v;
    asm("v_mov_b64 %0, %1" : "=v"(v) : "s"(s));
    return v;
#else
    return s;
#endif
}

template <int DT>
__device__ inline LdsRun lds_run64(const double *gp, double fi, double fk, double fj0,
                                   float scale, float view_scale, int u0, int v0,
                                   unsigned buf_byte, int cmax)
{
    const __attribute__((address_space(4))) double *g = as_constant(gp);
    const double den = fma(g[7], fi, fma(g[8], fk, uniform_to_vgpr64(g[9])));
    double r = __builtin_amdgcn_rcp(den);
    r = fma(r, fma(-den, r, 1.0), r);
    const double su =
        fma(fma(g[0], fi, fma(g[1], fk, uniform_to_vgpr64(g[2]))), r, (double)(1 - u0));
    const double sv =
        fma(fma(g[3], fj0, fma(g[4], fi, fma(g[5], fk, uniform_to_vgpr64(g[6])))), r,
            (double)((row_biased<DT>() ? 129 : 1) - v0));
    int col;
    float wu;
    if constexpr (DT == CTR_F16) {
        const float s = (float)su;
        const float fl = floorf(s);
        wu = s - fl;
        col = (int)fl;
    } else {
        const double fl = floor(su);
        wu = (float)(su - fl);
        col = (int)fl;
    }
    LdsRun R;
    // the two wave-uniform factors applied one after the other, so that
    // each multiply takes its factor straight from an SGPR
    const float w2 = (float)(r * r) * scale * view_scale;
    R.cR = w2 * wu;
    R.cL = w2 - R.cR;
    R.f = (float)sv;
    R.dv = (float)(g[3] * r);
    R.base = buf_byte + (unsigned)min(max(col, 0), cmax - 2) * 4u;
    return R;
}


template <int DT, int TY>
__device__ inline void lds_run_accumulate(float (&acc)[TY], const LdsRun &R, const unsigned *lds,
                                          unsigned lp4)
{
    // Software-pipelined in batches of kB voxels: the taps of batch b + 1
    // are in flight while batch b is interpolated; the scheduling barriers
    // keep that order.
    const unsigned base = lds_base<DT>(R.base, lp4);
    if constexpr (DT == CTR_F32) {
        constexpr int kB = CTR_F32_BATCH;
        float2 ta[2][kB], tb[2][kB];
        float tw[2][kB];
        auto issue = [&](int set, int t0) {
#pragma unroll
            for (int q = 0; q < kB; ++q) {
                // the first voxel's row is f itself (the FMA with 0 is not
                // folded away by the compiler)
                const float x = t0 + q == 0 ? R.f : fmaf(R.dv, (float)(t0 + q), R.f);
                tw[set][q] = __builtin_amdgcn_fractf(x);
                lds_taps_f32(lds, lds_row_addr<DT>(x, lp4, base), ta[set][q], tb[set][q]);
            }
        };
        issue(0, 0);
#pragma unroll
        for (int bt = 0; bt < TY / kB; ++bt) {
            if (bt + 1 < TY / kB) issue((bt + 1) & 1, (bt + 1) * kB);
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int q = 0; q < kB; ++q) {
                const float2 a = ta[bt & 1][q], b = tb[bt & 1][q];
                const float wv = tw[bt & 1][q];
                const float left = fmaf(wv, b.x - a.x, a.x);
                const float right = fmaf(wv, b.y - a.y, a.y);
                acc[bt * kB + q] = fmaf(R.cR, right, fmaf(R.cL, left, acc[bt * kB + q]));
            }
            __builtin_amdgcn_sched_barrier(0);
        }
    } else {
        // 16-bit value-and-step texels: columns u and u + 1 of row I (one
        // ds_read2_b32 of neighbouring dwords); each column's vertical lerp
        // v + wv d is one FMA on the texel, then the two weighted
        // accumulations
        static_assert(DT == CTR_F16 || DT == CTR_BF16, "staged format");
        constexpr int kB = CTR_16_BATCH;
        unsigned ta[2][kB], tb[2][kB];
        float tw[2][kB];
        auto issue = [&](int set, int t0) {
#pragma unroll
            for (int q = 0; q < kB; ++q) {
                // the first voxel's row is f itself (the FMA with 0 is not
                // folded away by the compiler)
                const float x = t0 + q == 0 ? R.f : fmaf(R.dv, (float)(t0 + q), R.f);
                tw[set][q] = __builtin_amdgcn_fractf(x);
                const unsigned *P = reinterpret_cast<const unsigned *>(
                    reinterpret_cast<const char *>(lds) + lds_row_addr<DT>(x, lp4, base));
                ta[set][q] = P[0];
                tb[set][q] = P[1];
            }
        };
        issue(0, 0);
#pragma unroll
        for (int bt = 0; bt < TY / kB; ++bt) {
            if (bt + 1 < TY / kB) issue((bt + 1) & 1, (bt + 1) * kB);
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int q = 0; q < kB; ++q) {
                const float wv = tw[bt & 1][q];
                const float left = vstep_lerp<DT>(ta[bt & 1][q], wv);
                const float right = vstep_lerp<DT>(tb[bt & 1][q], wv);
                float &A = acc[bt * kB + q];
                A = fmaf(R.cR, right, fmaf(R.cL, left, A));
            }
            __builtin_amdgcn_sched_barrier(0);
        }
    }
}

// First view a workgroup visits: 0, so every tile accumulates its views in
// the same order and results are reproducible bit for bit.
__device__ inline int first_view(int n_views)
{
    (void)n_views;
    return 0;
}

template <int DT, int TY>
__device__ inline void backproject_lds_tile(const BpJob &J, int bx, int by, int bz,
                                            unsigned *lds)
{
    using L = LdsLayout<DT>;
    constexpr int HMAX = L::kRows, CMAX = L::kCols;
    static_assert(HMAX < 127, "biased LDS rows hold rows below 128");
    constexpr int kBelow = DT == CTR_F32 ? 1 : 0;  // taps of row I + 1 read from LDS
    int4 *rects = reinterpret_cast<int4 *>(lds);
    unsigned *bufs = lds + L::kRectBytes / 4;
    const int pitch_u = J.pitch_u, pitch_v = J.pitch_v, N = J.n_views;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    constexpr int TW = L::kTile / 8;  // waves along x
    const int i = bx * L::kTile + (wave % TW) * 8 + (lane & 7);
    const int k = by * L::kTile + (wave / TW) * 8 + (lane >> 3);
    const int j0 = bz * TY;
    // the run's rows relative to the volume centre, and the staged v centre
    const float oy = 0.5f - 0.5f * (float)J.Y;
    const float ja = (float)j0 + oy, jb = (float)(j0 + TY - 1) + oy;
    const float cv1 = 0.5f * (float)J.nv + 0.5f;
    const float4 *foot = J.foot + (size_t)by * J.gx + bx;
    const size_t foot_stride = (size_t)J.gx * J.gy;
    const char *vbase = reinterpret_cast<const char *>(J.staged);
    const size_t view_bytes = J.view_texels * 4;
    const unsigned view_bytes32 = (unsigned)view_bytes;  // < 2^31, checked by the host
    const unsigned lp4 = J.lds_pitch4;

    // views are visited from v0 on, wrapping (v0 = 0: in order)
    const int v0 = first_view(N);
    auto vw = [&](int pos) {
        const int v = pos + v0;
        return v >= N ? v - N : v;
    };

    // clear the row buffers (every LDS tap then reads a finite value), fill
    // the first rectangle table
    for (int t = threadIdx.x; t < L::kBufDwords / 2; t += L::kThreads)
        reinterpret_cast<uint4 *>(bufs)[t] = make_uint4(0u, 0u, 0u, 0u);
    if (threadIdx.x <= kRectChunk && (int)threadIdx.x < N)
        rects[threadIdx.x] = tile_rect(J, foot, foot_stride, vw(threadIdx.x), ja, jb, cv1, HMAX, CMAX, kBelow);
    __syncthreads();

    float acc[TY];
#pragma unroll
    for (int t = 0; t < TY; ++t) acc[t] = 0.0f;

    int4 cur = read_rect(rects, 0);
    if (rect_mode(cur) == RECT_LDS)
        stage_rect<L::kPitch, CMAX, L::kWaves>(vbase + (size_t)vw(0) * view_bytes, view_bytes32,
                                               pitch_u, pitch_v,
                              cur.x, cur.y, cur.z, rect_cols(cur), bufs, wave, lane);
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();

    bool any_direct = false;
    int n0 = 0;
#pragma unroll 1
    for (int n = 0; n < N; ++n) {
        const int c = n - n0;
        int4 nxt = make_int4(0, 0, 0, RECT_SKIP);
        if (n + 1 < N) {
            nxt = read_rect(rects, c + 1);
            if (rect_mode(nxt) == RECT_LDS)
                stage_rect<L::kPitch, CMAX, L::kWaves>(vbase + (size_t)vw(n + 1) * view_bytes,
                                                       view_bytes32,
                                      pitch_u, pitch_v, nxt.x, nxt.y, nxt.z, rect_cols(nxt),
                                      bufs + ((n + 1) & 1) * L::kBufDwords, wave, lane);
        }
        if (rect_mode(cur) == RECT_LDS) {
            const int view = vw(n);
            const unsigned buf_byte = (unsigned)(L::kRectBytes + (n & 1) * L::kBufDwords * 4);
#if CTR_FOLD_ORIGIN
            const float vsv = J.view_scale ? as_constant(J.view_scale)[view] : 1.0f;
            const LdsRun R = lds_run64<DT>(J.geom64 + (size_t)kGeom64 * view, i, k, j0, J.scale,
                                           vsv, cur.x, cur.y, buf_byte, CMAX);
#else
            float vs = J.scale;
            if (J.view_scale) vs *= as_constant(J.view_scale)[view];
            const ColumnView cv = column_view64(J.geom64 + (size_t)kGeom64 * view, i, k, j0, vs);
            const LdsRun R = lds_run<DT>(cv, cur.x, cur.y, buf_byte, CMAX);
#endif
            lds_run_accumulate<DT, TY>(acc, R, lds, lp4);
        }
        any_direct |= rect_mode(cur) == RECT_DIRECT;
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        __syncthreads();
        if (c + 1 == kRectChunk && n + 1 < N) {
            n0 += kRectChunk;
            if (threadIdx.x <= kRectChunk && n0 + (int)threadIdx.x < N)
                rects[threadIdx.x] =
                    tile_rect(J, foot, foot_stride, vw(n0 + threadIdx.x), ja, jb, cv1, HMAX, CMAX, kBelow);
            __syncthreads();
        }
        cur = nxt;
    }

    // views whose footprint exceeds the buffers: direct taps, clamped
    if (any_direct) {
        const unsigned pitch4 = (unsigned)pitch_u * 4u;
        const double su_hi = (double)J.nu + 1.0;
#pragma unroll 1
        for (int n = 0; n < N; ++n) {
            const int mode = __builtin_amdgcn_readfirstlane(
                rect_mode(tile_rect(J, foot, foot_stride, n, ja, jb, cv1, HMAX, CMAX, kBelow)));
            if (mode != RECT_DIRECT) continue;
            float vs = J.scale;
            if (J.view_scale) vs *= J.view_scale[n];
            const ColumnView cv = column_view64(J.geom64 + (size_t)kGeom64 * n, i, k, j0, vs);
            const Run R = make_run(cv, su_hi, 0.0, 0, J.nv + 1, 0, pitch4, su_hi);
            const char *top = vbase + (size_t)n * view_bytes;
            gather_run<DT, TY>(acc, R, top, pitch4);
        }
    }
    store_run<TY>(acc, J.vol, i, j0, k, J.X, J.Y, J.Z, J.accumulate, J.clip);
}

// Tile order.  Workgroups reach the 8 XCDs round-robin (block b on XCD
// b mod 8).  bp_tile_order deals the tiles to the XCDs in chunks of
// kXcdChunk consecutive tiles, chunk c to XCD c mod 8, so that the XCDs work
// on neighbouring chunks of the same y-run layer at the same time and read
// the same detector rows (with one contiguous range of tiles per XCD, each
// XCD read rows of its own).  Inside a layer, lds_tile_coords walks bands of kTileBand tile
// rows along z, column by column, so that tiles dispatched together form a
// compact patch of columns whose rows they share in L2.
//
// Measured (FP16 / FP32 kernel ms):
//   512^3 x 720: rows along x, one contiguous range per XCD 25.1 / 33.2;
//     bands of 4: 23.3-23.6 / 31.7-32.3 (bands of 3-6 alike, 2, 8, 16
//     slower; 4 x 4 squares, serpentine or x bands no better); with chunks
//     of 128 tiles dealt round-robin 22.7-23.0 / 29.4-30.0 (chunks of 32, 64,
//     256 and 512 between the two).  Bytes fetched from beyond L2 per FP16
//     launch: 70.3, 43.8 and 31.2 GB.
//   1024^3 x 1200: 287.6 / 362.1, 274.4-279.7 / 349.4-354.1, 275.7-276.2 /
//     350.1-355.0.
//   1440^3 x 1440: bands of 4 in ranges 965 / 1159; chunks of 128 828 / 1014.
//   With chunks dealt round-robin, against bands of 4 in chunks of 128
//   (512^3 / 1024^3 FP16 kernel, alternating runs on one box): bands of 2
//   +1.5% / +0.8%, of 8 +0.7% / +2.1%, of 16 +2.8% / +2.2%; bands of 8 in
//   chunks of 64 +1.5% / +1.9%, of 256 +0.4% / +1.1% (FP32 +1.7-4.6% at
//   512^3).
#ifndef CTR_TILE_BAND
#define CTR_TILE_BAND 4
#endif
constexpr int kTileBand = CTR_TILE_BAND;
#ifndef CTR_XCD_CHUNK
#define CTR_XCD_CHUNK 128
#endif
constexpr int kXcdChunk = CTR_XCD_CHUNK;

// A bijection of [0, nblocks): XCD x's k-th block takes tile
// ((k / C) * 8 + x) * C + k mod C; the tail past the last whole round of
// 8 chunks keeps its own index.
__device__ inline int bp_tile_order(int b, int nblocks)
{
    constexpr int C = kXcdChunk;
    const int full = nblocks / (8 * C) * (8 * C);
    if (b >= full) return b;
    const int x = b & 7, k = b >> 3;
    return ((k / C) * 8 + x) * C + k % C;
}

__device__ inline void lds_tile_coords(int t, const BpJob &J, int &bx, int &by, int &bz)
{
    const int layer = J.gx * J.gy;
    bz = t / layer;
    const int r = t - bz * layer;
    const int band = r / (J.gx * kTileBand);
    const int h = min(kTileBand, J.gy - band * kTileBand);
    const int rb = r - band * J.gx * kTileBand;
    bx = rb / h;
    by = band * kTileBand + (rb - bx * h);
}

// Waves per SIMD the compiler must leave registers for (the second launch
// bound; one wave of each workgroup per SIMD, so also workgroups per CU), fp32
// and 16-bit tiles.  1 leaves the register budget to the compiler.  Measured
// (GUPS, 512^3 / 1024^3 / 1440^3): fp32 at 5 (74 VGPRs in place of 120) with
// 70 rows of pitch 53 2920 / 3708 / 3610, with 72 rows of pitch 51 2877 /
// 3685 / 3697, against 2950 / 3621 / 3763 at 4 workgroups (84 rows): the
// taller footprints that then take the direct pass cost more than the fifth
// workgroup gains.  16-bit at 6 (64-68 VGPRs, 60 rows) 2564 / 3840 / 2577
// against 3414 / 3843 / 3956 for fp16: no gain even where every footprint
// fits (1024^3).
#ifndef CTR_F32_MINB
#define CTR_F32_MINB 1
#endif
#ifndef CTR_16_MINB
#define CTR_16_MINB 1
#endif
template <int DT>
constexpr int lds_min_blocks()
{
    return DT == CTR_F32 ? CTR_F32_MINB : CTR_16_MINB;
}

template <int DT, int TY>
__global__ __launch_bounds__(LdsLayout<DT>::kThreads, lds_min_blocks<DT>()) void backproject_lds_kernel(
    BpJob J)
{
    __shared__ __align__(16) unsigned lds[LdsLayout<DT>::kDwords];
    int bx, by, bz;
    lds_tile_coords(bp_tile_order(blockIdx.x, gridDim.x), J, bx, by, bz);
    backproject_lds_tile<DT, TY>(J, bx, by, bz, lds);
}

// Packed launch: the flat tile index selects the job (the last job whose
// block_begin it reaches) and the tile inside it.
template <int DT, int TY>
__global__ __launch_bounds__(LdsLayout<DT>::kThreads, lds_min_blocks<DT>()) void backproject_lds_packed_kernel(
    BpBatch B)
{
    __shared__ __align__(16) unsigned lds[LdsLayout<DT>::kDwords];
    const int b = bp_tile_order(blockIdx.x, gridDim.x);
    int j = 0;
#pragma unroll 1
    for (int q = 1; q < B.count; ++q)
        if (b >= B.job[q].block_begin) j = q;
    const BpJob &J = B.job[j];
    int bx, by, bz;
    lds_tile_coords(b - J.block_begin, J, bx, by, bz);
    backproject_lds_tile<DT, TY>(J, bx, by, bz, lds);
}

constexpr int kTY = CTR_LDS_TY;  // LDS kernel: voxels per thread along y

// Byte offsets inside one staged view are 32-bit: the row DMA's buffer
// resource (num_records a positive int) and the direct taps'
// mad24(row, pitch) + base, whose operands (row < 2^21, pitch < 2^24 bytes
// with nu < 2^20) fit 24 bits and whose 32-bit sum is exact modulo 2^32.
// A 2048^2 panel is 2080 x 2051 staged texels (17 MB).
constexpr size_t kMaxViewBytes = (size_t)1 << 31;

// Column tile edge of the LDS kernel for a staged format.
static int lds_tile(int staged_dtype)
{
    return staged_dtype == CTR_F32 ? LdsConfig<CTR_F32>::kTile : LdsConfig<CTR_F16>::kTile;
}

// `tile`: the column tile edge of the kernel that runs (lds_tile for the LDS
// kernel, 16 for the direct-tap kernel).
static BpJob make_job(const void *staged, const StagedLayout &L, int nu, int nv, GeomPtrs geom,
                      const float *view_scale, int n_views, float *vol, int X, int Y, int Z,
                      float scale, int accumulate, int clip, int tile)
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
    J.gx = (X + tile - 1) / tile;
    J.gy = (Z + tile - 1) / tile;
    J.foot = nullptr;
    J.block_begin = 0;
    J.lds_pitch4 = 0;  // set per staged format at launch
    return J;
}

// Kernel choice: the LDS-staged kernel for volumes at least 32 voxels along
// the rotation axis, the direct-tap kernel otherwise; the workspace's
// CTR_OPTION_BACKPROJECT_KERNEL overrides (diagnostics and A/B timing).
static bool use_lds_kernel(int kernel_option, int Y)
{
    return kernel_option == CTR_KERNEL_LDS || (kernel_option == CTR_KERNEL_AUTO && Y >= 32);
}

// Fills the footprint table of job J at `dst` (gx * gy * n_views float4)
// for tiles of `tile` x `tile` columns.
static int build_footprints(hipStream_t s, BpJob &J, int tile, float4 *dst)
{
    const long long total = (long long)J.gx * J.gy * J.n_views;
    footprint_kernel<<<(unsigned)((total + 255) / 256), 256, 0, s>>>(J.geom64, J.n_views, J.gx,
                                                                      J.gy, tile, dst);
    CTR_CHECK_LAUNCH();
    J.foot = dst;
    return CTR_OK;
}

static size_t footprint_bytes(int X, int Z, int n_views, int tile)
{
    return sizeof(float4) * (size_t)((X + tile - 1) / tile) * ((Z + tile - 1) / tile) * n_views;
}

template <int DT>
static void launch_lds(hipStream_t s, BpJob J, unsigned blocks)
{
    J.lds_pitch4 = LdsConfig<DT>::kPitch * 4u;
    backproject_lds_kernel<DT, kTY><<<blocks, LdsLayout<DT>::kThreads, 0, s>>>(J);
}

template <int DT>
static void launch_lds_packed(hipStream_t s, BpBatch B, unsigned blocks)
{
    for (int q = 0; q < B.count; ++q) B.job[q].lds_pitch4 = LdsConfig<DT>::kPitch * 4u;
    backproject_lds_packed_kernel<DT, kTY><<<blocks, LdsLayout<DT>::kThreads, 0, s>>>(B);
}

template <int DT>
static int launch_bp(hipStream_t s, const BpJob &J, bool lds)
{
    if (lds) {
        const long long blocks = (long long)J.gx * J.gy * ((J.Y + kTY - 1) / kTY);
        CTR_REQUIRE(blocks < (1LL << 31), CTR_ERR_INVALID_ARGUMENT, "volume too large");
        launch_lds<DT>(s, J, (unsigned)blocks);
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

// Host-side checks shared by ctr_backproject and ctr_backproject_noalloc.
static int check_backproject(int geometry, const void *staged, int staged_dtype,
                             const float *view_scale, int nu, int nv, int n_views, double sod,
                             double odd, double detector_width, double volume_width,
                             const float *volume, int X, int Y, int Z)
{
    CTR_REQUIRE(staged && volume, CTR_ERR_NULL_POINTER,
                "[CTAnalyzer Error] ctr_backproject: projections or volume is NULL");
    CTR_REQUIRE(staged_dtype_valid(staged_dtype), CTR_ERR_INVALID_ARGUMENT,
                "unknown staged dtype %d", staged_dtype);
    CTR_REQUIRE(staged_dtype == CTR_F32 || view_scale, CTR_ERR_NULL_POINTER,
                "view_scale is required for 16-bit staged projections");
    int rc = check_geometry(geometry, sod, odd, detector_width, volume_width, X, Y, Z, nu, nv,
                            n_views);
    if (rc) return rc;
    StagedLayout L = staged_layout(nu, nv, n_views, staged_dtype);
    CTR_REQUIRE(L.view * 4 < kMaxViewBytes, CTR_ERR_INVALID_ARGUMENT,
                "staged view of %zu bytes exceeds the backprojector's 32-bit in-view "
                "offsets (%zu bytes)", L.view * 4, kMaxViewBytes);
    return CTR_OK;
}

// One backprojection with its scratch given: the per-view geometry records
// at `geom_buf` (kGeomBytesPerView per view) and, when the LDS kernel runs,
// the tile footprint table at `foot_buf` (footprint_bytes).  Both tables are
// written whole before they are read, so no scratch needs clearing.
static int backproject_with(hipStream_t s, int kernel_option, int geometry,
                            const void *staged, int staged_dtype, const float *view_scale,
                            int nu, int nv, int n_views, double sod, double odd,
                            double detector_width, double volume_width, double rotation_count,
                            double rotation_offset, float *volume, int X, int Y, int Z,
                            float scale, int accumulate, int clip_negative, void *geom_buf,
                            void *foot_buf)
{
    GeomPtrs g;
    int rc = build_geometry_at(s, geom_buf, geometry, n_views, sod, odd, detector_width,
                               volume_width, rotation_count, rotation_offset, X, Y, Z, nu, nv, &g);
    if (rc) return rc;
    StagedLayout L = staged_layout(nu, nv, n_views, staged_dtype);
    const bool lds = use_lds_kernel(kernel_option, Y);
    const int tile = lds ? lds_tile(staged_dtype) : 16;
    BpJob J = make_job(staged, L, nu, nv, g, view_scale, n_views, volume, X, Y, Z, scale,
                       accumulate, clip_negative, tile);
    if (lds && (rc = build_footprints(s, J, tile, (float4 *)foot_buf))) return rc;
    switch (staged_dtype) {
    case CTR_F32: return launch_bp<CTR_F32>(s, J, lds);
    case CTR_F16: return launch_bp<CTR_F16>(s, J, lds);
    default: return launch_bp<CTR_BF16>(s, J, lds);
    }
}

// Scratch of ctr_backproject_noalloc: the geometry and footprint tables in a
// device array of the library itself (each GPU holds its own copy for the
// life of the process), so that entry point allocates nothing.  Host code
// uses the per-device address from hipGetSymbolAddress, never the array's
// name.  Every ctr_backproject_noalloc call on a GPU shares its array:
// calls that overlap on one GPU race.
constexpr size_t kNoAllocScratchBytes = (size_t)128 << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];
static std::mutex g_scratch_lock;
static void *g_scratch_ptr[64] = {nullptr};

// The calling device's address of `scratch`, looked up once per device.
static int scratch_address(void **out)
{
    int dev = 0;
    CTR_TRY_HIP(hipGetDevice(&dev));
    CTR_REQUIRE(dev >= 0 && dev < 64, CTR_ERR_INVALID_ARGUMENT, "device index %d out of range",
                dev);
    std::lock_guard<std::mutex> g(g_scratch_lock);
    if (!g_scratch_ptr[dev]) {
        void *ptr = nullptr;
        CTR_TRY_HIP(hipGetSymbolAddress((void **)&ptr, HIP_SYMBOL(scratch)));
        CTR_REQUIRE(ptr, CTR_ERR_HIP, "hipGetSymbolAddress(scratch) returned NULL");
        g_scratch_ptr[dev] = ptr;
    }
    *out = g_scratch_ptr[dev];
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
    int rc = check_backproject(geometry, staged, staged_dtype, view_scale, nu, nv, n_views,
                               source_distance, detector_distance, detector_width, volume_width,
                               volume, X, Y, Z);
    if (rc) return rc;
    ctr_workspace *w = nullptr;
    if ((rc = workspace_resolve(ws, &w))) return rc;
    const int kernel_option = workspace_option(w, CTR_OPTION_BACKPROJECT_KERNEL);
    void *geom = nullptr, *foot = nullptr;
    if ((rc = workspace_get(w, SLOT_GEOMETRY, kGeomBytesPerView * (size_t)n_views, &geom)))
        return rc;
    if (use_lds_kernel(kernel_option, Y) &&
        (rc = workspace_get(w, SLOT_FOOTPRINT,
                            footprint_bytes(X, Z, n_views, lds_tile(staged_dtype)), &foot)))
        return rc;
    return backproject_with(as_stream(stream), kernel_option, geometry, staged,
                            staged_dtype, view_scale, nu, nv, n_views, source_distance,
                            detector_distance, detector_width, volume_width, rotation_count,
                            rotation_offset, volume, X, Y, Z, scale, accumulate, clip_negative,
                            geom, foot);
}

int ctr_backproject_noalloc(void *stream, int geometry, const void *staged, int staged_dtype,
                            const float *view_scale, int nu, int nv, int n_views,
                            double source_distance, double detector_distance,
                            double detector_width, double volume_width, double rotation_count,
                            double rotation_offset, float *volume, int X, int Y, int Z,
                            float scale, int accumulate, int clip_negative)
{
    clear_error();
    int rc = check_backproject(geometry, staged, staged_dtype, view_scale, nu, nv, n_views,
                               source_distance, detector_distance, detector_width, volume_width,
                               volume, X, Y, Z);
    if (rc) return rc;
    const int kernel_option = default_option(CTR_OPTION_BACKPROJECT_KERNEL);
    const size_t geom_bytes = (kGeomBytesPerView * (size_t)n_views + 255) & ~(size_t)255;
    const int tile = lds_tile(staged_dtype);
    const size_t foot =
        use_lds_kernel(kernel_option, Y) ? footprint_bytes(X, Z, n_views, tile) : 0;
    CTR_REQUIRE(geom_bytes + foot <= kNoAllocScratchBytes, CTR_ERR_OVER_BUDGET,
                "ctr_backproject_noalloc: %d views over a %d x %d tile grid need %zu bytes of "
                "scratch, more than its fixed %zu (use ctr_backproject)", n_views,
                (X + tile - 1) / tile, (Z + tile - 1) / tile, geom_bytes + foot,
                kNoAllocScratchBytes);
    void *base = nullptr;
    if ((rc = scratch_address(&base))) return rc;
    return backproject_with(as_stream(stream), kernel_option, geometry, staged,
                            staged_dtype, view_scale, nu, nv, n_views, source_distance,
                            detector_distance, detector_width, volume_width, rotation_count,
                            rotation_offset, volume, X, Y, Z, scale, accumulate, clip_negative,
                            base, (char *)base + geom_bytes);
}

int ctr_backproject_packed(void *stream, ctr_workspace *ws, int staged_dtype,
                           const ctr_backproject_job *jobs, int count)
{
    clear_error();
    CTR_REQUIRE(jobs, CTR_ERR_NULL_POINTER, "ctr_backproject_packed: jobs is NULL");
    CTR_REQUIRE(count >= 1 && count <= kMaxPacked, CTR_ERR_INVALID_ARGUMENT,
                "ctr_backproject_packed: count %d outside [1, %d]", count, kMaxPacked);
    CTR_REQUIRE(staged_dtype_valid(staged_dtype), CTR_ERR_INVALID_ARGUMENT,
                "unknown staged dtype %d", staged_dtype);
    size_t total_views = 0, total_foot = 0;
    for (int q = 0; q < count; q++) {
        const ctr_backproject_job &j = jobs[q];
        total_foot += footprint_bytes(j.X, j.Z, j.n_views, lds_tile(staged_dtype));
        CTR_REQUIRE(j.staged && j.volume, CTR_ERR_NULL_POINTER,
                    "ctr_backproject_packed: job %d has a NULL buffer", q);
        CTR_REQUIRE(staged_dtype == CTR_F32 || j.view_scale, CTR_ERR_NULL_POINTER,
                    "ctr_backproject_packed: job %d needs view_scale", q);
        int rc = check_geometry(j.geometry, j.source_distance, j.detector_distance,
                                j.detector_width, j.volume_width, j.X, j.Y, j.Z, j.nu, j.nv,
                                j.n_views);
        if (rc) return rc;
        StagedLayout L = staged_layout(j.nu, j.nv, j.n_views, staged_dtype);
        CTR_REQUIRE(L.view * 4 < kMaxViewBytes, CTR_ERR_INVALID_ARGUMENT,
                    "job %d: staged view of %zu bytes exceeds the backprojector's 32-bit "
                    "in-view offsets", q, L.view * 4);
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
        const int tile = lds_tile(staged_dtype);
        B.job[q] = make_job(j.staged, L, j.nu, j.nv, gp, j.view_scale, j.n_views, j.volume, j.X,
                            j.Y, j.Z, j.scale, j.accumulate, j.clip_negative, tile);
        if ((rc = build_footprints(s, B.job[q], tile, (float4 *)fp))) return rc;
        B.job[q].block_begin = (int)blocks;
        blocks += (long long)B.job[q].gx * B.job[q].gy * ((j.Y + kTY - 1) / kTY);
        g += kGeomBytesPerView * (size_t)j.n_views;
        fp += footprint_bytes(j.X, j.Z, j.n_views, tile);
    }
    CTR_REQUIRE(blocks < (1LL << 31), CTR_ERR_INVALID_ARGUMENT, "packed launch too large");
    switch (staged_dtype) {
    case CTR_F32: launch_lds_packed<CTR_F32>(s, B, (unsigned)blocks); break;
    case CTR_F16: launch_lds_packed<CTR_F16>(s, B, (unsigned)blocks); break;
    default: launch_lds_packed<CTR_BF16>(s, B, (unsigned)blocks); break;
    }
    CTR_CHECK_LAUNCH();
    return CTR_OK;
}

}  // extern "C"
