// forward.hip -- ray-driven forward projection, fan->parallel rebinning and
// the elementwise detector / volume passes, for gfx950.
#include <cmath>
#include <vector>

#include "ctr_internal.h"

namespace ctr {

static inline unsigned grid1d(long long total, int block)
{
    return (unsigned)((total + block - 1) / block);
}

// ---------------------------------------------------------------------------
// Forward projection
// ---------------------------------------------------------------------------
__device__ inline float vol_at(const float *__restrict__ vol, int X, int Y, int Z, int i, int j,
                               int k)
{
    return (i >= 0 && i < X && j >= 0 && j < Y && k >= 0 && k < Z)
               ? vol[((size_t)k * Y + j) * X + i]
               : 0.0f;
}

// Trilinear sample at continuous voxel-index coordinates, zero outside.
__device__ inline float trilinear0(const float *__restrict__ vol, int X, int Y, int Z, float fi,
                                   float fj, float fk)
{
    if (!(fi > -1.0f && fi < (float)X && fj > -1.0f && fj < (float)Y && fk > -1.0f &&
          fk < (float)Z))
        return 0.0f;
    const float i0f = floorf(fi), j0f = floorf(fj), k0f = floorf(fk);
    const int i0 = (int)i0f, j0 = (int)j0f, k0 = (int)k0f;
    const float wi = fi - i0f, wj = fj - j0f, wk = fk - k0f;
    float c00, c01, c10, c11;
    if (i0 >= 0 && i0 + 1 < X && j0 >= 0 && j0 + 1 < Y && k0 >= 0 && k0 + 1 < Z) {
        const float *p = vol + ((size_t)k0 * Y + j0) * X + i0;
        const size_t sy = X, sz = (size_t)X * Y;
        c00 = fmaf(wi, p[1] - p[0], p[0]);
        c01 = fmaf(wi, p[sy + 1] - p[sy], p[sy]);
        c10 = fmaf(wi, p[sz + 1] - p[sz], p[sz]);
        c11 = fmaf(wi, p[sz + sy + 1] - p[sz + sy], p[sz + sy]);
    } else {
        float a, b;
        a = vol_at(vol, X, Y, Z, i0, j0, k0);
        b = vol_at(vol, X, Y, Z, i0 + 1, j0, k0);
        c00 = fmaf(wi, b - a, a);
        a = vol_at(vol, X, Y, Z, i0, j0 + 1, k0);
        b = vol_at(vol, X, Y, Z, i0 + 1, j0 + 1, k0);
        c01 = fmaf(wi, b - a, a);
        a = vol_at(vol, X, Y, Z, i0, j0, k0 + 1);
        b = vol_at(vol, X, Y, Z, i0 + 1, j0, k0 + 1);
        c10 = fmaf(wi, b - a, a);
        a = vol_at(vol, X, Y, Z, i0, j0 + 1, k0 + 1);
        b = vol_at(vol, X, Y, Z, i0 + 1, j0 + 1, k0 + 1);
        c11 = fmaf(wi, b - a, a);
    }
    const float c0 = fmaf(wj, c01 - c00, c00);
    const float c1 = fmaf(wj, c11 - c10, c10);
    return fmaf(wk, c1 - c0, c0);
}

struct ForwardJob {
    int geometry;
    const float *vol;
    int X, Y, Z;
    double d, sod, odd, delta, rot, offset, h;
    float *out;
    int layout, nu, nv, n_views;
};

// Samples k of a ray whose index-space point f(k) = b + (k + 1/2) s lies in
// [m, n - 1 - m] on every axis (n the axis length): a contiguous range
// [*ka, *kb) of [0, K), empty when *ka >= *kb.  Such a sample's trilinear
// cell (i0, i0 + 1) is inside the volume on every axis.  b and s are the
// FP32 values the march uses; the margin m = 1/64 covers the one rounding
// of fmaf((float)k + 0.5f, s, b) for axes up to 2^17 voxels.
__device__ inline void interior_samples(const float b[3], const float s[3], const int n[3],
                                        int K, int *ka, int *kb)
{
    const double m = 1.0 / 64.0;
    double kl = 0.0, kh = (double)K - 1.0;
    for (int a = 0; a < 3; ++a) {
        const double f_lo = m, f_hi = (double)n[a] - 1.0 - m;
        if (f_hi < f_lo) {
            kh = -1.0;
            break;
        }
        if (s[a] == 0.0f) {
            if ((double)b[a] < f_lo || (double)b[a] > f_hi) kh = -1.0;
            continue;
        }
        double t0 = (f_lo - (double)b[a]) / (double)s[a];
        double t1 = (f_hi - (double)b[a]) / (double)s[a];
        if (t0 > t1) {
            const double t = t0;
            t0 = t1;
            t1 = t;
        }
        kl = fmax(kl, ceil(t0 - 0.5));
        kh = fmin(kh, floor(t1 - 0.5));
    }
    if (kh < kl) {
        *ka = 0;
        *kb = 0;
        return;
    }
    *ka = (int)kl;
    *kb = (int)kh + 1;
}

// Trilinear sample of an interior point (its cell inside the volume on every
// axis, f > 0): the arithmetic of trilinear0's in-volume branch, with the
// cell's byte offset in 32 bits and the four rows of the cell read through a
// buffer resource at SGPR offsets 0, 4X, 4XY, 4XY + 4X.  Needs the volume
// below 4 GiB and X * Y below 2^24.
__device__ inline float trilinear_interior_buf(__amdgpu_buffer_rsrc_t vol, unsigned X,
                                               unsigned XY, unsigned sy4, unsigned sz4,
                                               float fi, float fj, float fk)
{
    const int i0 = (int)fi, j0 = (int)fj, k0 = (int)fk;
    const float wi = fi - (float)i0, wj = fj - (float)j0, wk = fk - (float)k0;
    const unsigned off =
        __umul24((unsigned)k0, XY) * 4u + __umul24((unsigned)j0, X) * 4u + (unsigned)i0 * 4u;
    auto ld = [&](unsigned o, unsigned so) {
        return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(vol, o, so, 0));
    };
    const float p0 = ld(off, 0), p1 = ld(off + 4, 0);
    const float q0 = ld(off, sy4), q1 = ld(off + 4, sy4);
    const float r0 = ld(off, sz4), r1 = ld(off + 4, sz4);
    const float t0 = ld(off, sz4 + sy4), t1 = ld(off + 4, sz4 + sy4);
    const float c00 = fmaf(wi, p1 - p0, p0);
    const float c01 = fmaf(wi, q1 - q0, q0);
    const float c10 = fmaf(wi, r1 - r0, r0);
    const float c11 = fmaf(wi, t1 - t0, t0);
    const float c0 = fmaf(wj, c01 - c00, c00);
    const float c1 = fmaf(wj, c11 - c10, c10);
    return fmaf(wk, c1 - c0, c0);
}

// The same for volumes of 4 GiB or more: 64-bit cell address.
__device__ inline float trilinear_interior_ptr(const float *__restrict__ vol, int X, int Y,
                                               float fi, float fj, float fk)
{
    const int i0 = (int)fi, j0 = (int)fj, k0 = (int)fk;
    const float wi = fi - (float)i0, wj = fj - (float)j0, wk = fk - (float)k0;
    const size_t sy = (size_t)X, sz = (size_t)X * Y;
    const float *p = vol + ((size_t)k0 * Y + j0) * X + i0;
    const float c00 = fmaf(wi, p[1] - p[0], p[0]);
    const float c01 = fmaf(wi, p[sy + 1] - p[sy], p[sy]);
    const float c10 = fmaf(wi, p[sz + 1] - p[sz], p[sz]);
    const float c11 = fmaf(wi, p[sz + sy + 1] - p[sz + sy], p[sz + sy]);
    const float c0 = fmaf(wj, c01 - c00, c00);
    const float c1 = fmaf(wj, c11 - c10, c10);
    return fmaf(wk, c1 - c0, c0);
}

// Interior samples per unrolled step.  Measured (ms, 512^3 / 256^3 / 1440^3
// inputs, all bit-identical): the checked loop alone 1225 / 58.9 / 23395;
// buffer path unrolled 1, 2, 4: 1163 / 52.9, 1173 / 54.2, 1173 / 54.1;
// 64-bit path (1440^3, 12 GB) unrolled 1, 2, 4: 23651, 20666, 24726.
// Block order: 0 walks the views in order with consecutive blocks on
// different XCDs (all XCDs work on one view at a time); 1 hands each XCD a
// contiguous range of views.  Measured (ms, 512^3 / 256^3 / 1440^3): 0 1164
// / 53.1 / 20642, 1 1272 / 68.5 / 26997 -- with every XCD on the same view
// the volume streams through the shared last-level cache once per view.
// Index-space ray of detector pixel (iu, iv) in view n: sample k < K sits at
// voxel-index coordinates b + (k + 1/2) s (x, y, z), the ray clipped to the
// volume box grown by one voxel per side; K = 0 when it misses the box.  Set
// up in FP64 (FBP3D.cpp:280-284 angle convention), stepped in FP32.
struct IndexRay {
    float b[3], s[3];
    int K;
};

// World-space line of detector pixel (iu, iv) in view n: origin o and unit
// direction dir (cone: from the source).
__device__ inline void pixel_line(const ForwardJob &J, int n, int iu, int iv, double o[3],
                                  double dir[3])
{
    double th = 2.0 * M_PI * J.rot * (double)n / (double)J.n_views + J.offset;
    double s, c;
    sincos(th, &s, &c);
    double u = (iu + 0.5 - 0.5 * J.nu) * J.delta;
    double v = (iv + 0.5 - 0.5 * J.nv) * J.delta;
    if (J.geometry == CTR_GEOMETRY_CONE) {
        o[0] = J.sod * c;
        o[1] = 0.0;
        o[2] = J.sod * s;
        dir[0] = -J.odd * c - u * s - o[0];
        dir[1] = v;
        dir[2] = -J.odd * s + u * c - o[2];
        double len = sqrt(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
        dir[0] /= len;
        dir[1] /= len;
        dir[2] /= len;
    } else {
        o[0] = -u * s;
        o[1] = v;
        o[2] = u * c;
        dir[0] = -c;
        dir[1] = 0.0;
        dir[2] = -s;
    }
}

// The same line in voxel-index coordinates, FP32: L[0..2] a point, L[3..5]
// the direction (per unit world length).
__device__ inline void index_line(const ForwardJob &J, int n, int iu, int iv, float L[6])
{
    double o[3], dir[3];
    pixel_line(J, n, iu, iv, o, dir);
    const double inv_d = 1.0 / J.d;
    L[0] = (float)(o[0] * inv_d + 0.5 * J.X - 0.5);
    L[1] = (float)(o[1] * inv_d + 0.5 * J.Y - 0.5);
    L[2] = (float)(o[2] * inv_d + 0.5 * J.Z - 0.5);
    L[3] = (float)(dir[0] * inv_d);
    L[4] = (float)(dir[1] * inv_d);
    L[5] = (float)(dir[2] * inv_d);
}

__device__ inline IndexRay index_ray(const ForwardJob &J, int n, int iu, int iv)
{
    IndexRay R;
    R.K = 0;
    double o[3], dir[3];
    pixel_line(J, n, iu, iv, o, dir);
    const double half[3] = {(0.5 * J.X + 1.0) * J.d, (0.5 * J.Y + 1.0) * J.d,
                            (0.5 * J.Z + 1.0) * J.d};
    double tmin = -1e300, tmax = 1e300;
    bool hit = true;
    for (int a = 0; a < 3; a++) {
        if (dir[a] == 0.0) {
            if (o[a] < -half[a] || o[a] > half[a]) hit = false;
            continue;
        }
        double t0 = (-half[a] - o[a]) / dir[a], t1 = (half[a] - o[a]) / dir[a];
        if (t0 > t1) {
            double t = t0;
            t0 = t1;
            t1 = t;
        }
        tmin = fmax(tmin, t0);
        tmax = fmin(tmax, t1);
    }
    for (int a = 0; a < 3; a++) R.b[a] = R.s[a] = 0.0f;
    if (hit && tmax > tmin) {
        R.K = (int)ceil((tmax - tmin) / J.h);
        const double inv_d = 1.0 / J.d;
        R.b[0] = (float)((o[0] + tmin * dir[0]) * inv_d + 0.5 * J.X - 0.5);
        R.b[1] = (float)((o[1] + tmin * dir[1]) * inv_d + 0.5 * J.Y - 0.5);
        R.b[2] = (float)((o[2] + tmin * dir[2]) * inv_d + 0.5 * J.Z - 0.5);
        R.s[0] = (float)(dir[0] * J.h * inv_d);
        R.s[1] = (float)(dir[1] * J.h * inv_d);
        R.s[2] = (float)(dir[2] * J.h * inv_d);
    }
    return R;
}

#ifndef CTR_FWD_XCD
#define CTR_FWD_XCD 0
#endif
// Detector pixels of a 256-thread block along u (the rest along v): compact
// tiles keep the block's rays close in the volume, so their gathers share
// cache lines.  Measured (ms, 512^3 / 256^3 inputs): 4 wide 1234 / 58.1,
// 8 wide 1184 / 52.0, 16 wide 1160 / 52.9, 32 wide 1284 / 58.6, 64 wide
// 1626 / 70.5.
#ifndef CTR_FWD_BU
#define CTR_FWD_BU 16
#endif
#ifndef CTR_FWD_UNROLL_BUF
#define CTR_FWD_UNROLL_BUF 1
#endif
#ifndef CTR_FWD_UNROLL_PTR
#define CTR_FWD_UNROLL_PTR 2
#endif

template <bool kBuf>
__global__ __launch_bounds__(256) void forward_kernel(ForwardJob J)
{
    // 1D grid of (view, v tile, u tile)
    const int tu = (J.nu + CTR_FWD_BU - 1) / CTR_FWD_BU;
    const int tv = (J.nv + 256 / CTR_FWD_BU - 1) / (256 / CTR_FWD_BU);
    const int b = CTR_FWD_XCD ? xcd_tile_order(blockIdx.x, gridDim.x) : (int)blockIdx.x;
    const int n = b / (tu * tv), tile = b % (tu * tv);
    const int iu = (tile % tu) * CTR_FWD_BU + (threadIdx.x % CTR_FWD_BU);
    const int iv = (tile / tu) * (256 / CTR_FWD_BU) + (threadIdx.x / CTR_FWD_BU);
    if (iu >= J.nu || iv >= J.nv) return;
    const IndexRay R = index_ray(J, n, iu, iv);
    float acc = 0.0f;
    if (R.K > 0) {
        const int K = R.K;
        const float bi = R.b[0], bj = R.b[1], bk = R.b[2];
        const float si = R.s[0], sj = R.s[1], sk = R.s[2];
        // samples in order: checked ones before the interior range, the
        // interior ones without bounds tests, checked ones after it (the sum
        // is the same sequence of additions as one checked loop)
        const float b3[3] = {bi, bj, bk}, s3[3] = {si, sj, sk};
        const int n3[3] = {J.X, J.Y, J.Z};
        int ka, kb;
        interior_samples(b3, s3, n3, K, &ka, &kb);
        if (kb <= ka) ka = kb = K;
        for (int k = 0; k < ka; ++k) {
            const float t = (float)k + 0.5f;
            acc += trilinear0(J.vol, J.X, J.Y, J.Z, fmaf(t, si, bi), fmaf(t, sj, bj),
                              fmaf(t, sk, bk));
        }
        if constexpr (kBuf) {
            const unsigned X = (unsigned)J.X, XY = (unsigned)J.X * (unsigned)J.Y;
            const __amdgpu_buffer_rsrc_t rsrc = __builtin_amdgcn_make_buffer_rsrc(
                const_cast<float *>(J.vol), (short)0,
                (int)(XY * (unsigned)J.Z * 4u), 0x00020000);
#pragma unroll CTR_FWD_UNROLL_BUF
            for (int k = ka; k < kb; ++k) {
                const float t = (float)k + 0.5f;
                acc += trilinear_interior_buf(rsrc, X, XY, X * 4u, XY * 4u, fmaf(t, si, bi),
                                              fmaf(t, sj, bj), fmaf(t, sk, bk));
            }
        } else {
#pragma unroll CTR_FWD_UNROLL_PTR
            for (int k = ka; k < kb; ++k) {
                const float t = (float)k + 0.5f;
                acc += trilinear_interior_ptr(J.vol, J.X, J.Y, fmaf(t, si, bi),
                                              fmaf(t, sj, bj), fmaf(t, sk, bk));
            }
        }
        for (int k = kb; k < K; ++k) {
            const float t = (float)k + 0.5f;
            acc += trilinear0(J.vol, J.X, J.Y, J.Z, fmaf(t, si, bi), fmaf(t, sj, bj),
                              fmaf(t, sk, bk));
        }
        acc *= (float)J.h;
    }
    size_t idx = J.layout == CTR_LAYOUT_PROJECTIONS
                     ? ((size_t)n * J.nv + iv) * J.nu + iu
                     : ((size_t)iv * J.n_views + n) * J.nu + iu;
    J.out[idx] = acc;
}

// ---------------------------------------------------------------------------
// Slab-staged forward projection.  The block's 256 rays (one kSlabTU x
// kSlabTV detector tile of one view) march together through slabs of kSlabS
// slices along the view's main horizontal axis A (x when |cos theta| >=
// |sin theta|, else z), consecutive slabs sharing one slice.  For each slab
// the block stages the box of voxels its rays' samples can touch there --
// zero outside the volume -- in LDS, and every ray takes the samples whose
// cell starts in the slab from LDS.  Each ray still sums its samples in order
// with trilinear0's arithmetic on the same taps, so the result equals
// forward_kernel's bit for bit.  A slab whose box exceeds the LDS region
// reads the volume directly, and a ray that does not run along +A (or -A)
// like the view's central ray, or a sample outside every slab, falls back to
// trilinear0 after the slabs, in order.
//
// The box is staged as [A][y][C], the cross axis C (z when A = x, x when
// A = z) fastest, and read from memory row by row along C: from the volume
// itself when C = x, from a copy with z fastest (volT, (x, y, z) at
// (x Y + y) Z + z) when C = z.  Each box row is then one contiguous run of
// Rc voxels in memory, and the detector tile's u side -- which spans C --
// sets its length.
// ---------------------------------------------------------------------------
// Slab depth, box capacity (floats) and the detector tile (u x v rays).
// Measured (ms; 512^3 x 720 views of 768^2, 256^3 x 360 of 384^2, 1440^3 x
// 96 of 1536^2; all bit-identical), this kernel: 8 slices / 24 KB 347 / 22.5
// / 1138; 6 / 18 KB 378 / 24.7 / 1161; 7 / 21 KB 360 / 23.5 / 1144; 8 /
// 20 KB 365 / 22.7 / 1202; 8 / 28 KB 377 / 24.4 / 1220; 9 / 27 KB 370 / 24.0
// / 1302; 10 / 30 KB 367 / 23.5 / 1318; 12 / 36 KB 412 / 26.2 / 1586; 32 x 8
// rays, 8 / 24 KB 393 / 24.1 / 1254.  With the earlier per-ray box bounds
// (wave reductions) and plain loads: 16 x 16, 6 / 18 KB 434 / 27.9 / 1144;
// 32 x 8 623, 64 x 4 845 (512^3, before the DMA fill).  The first slab
// kernel (box x fastest, read along x for every view, element indices by
// integer division) measured 667 / 41.2 / 1393.
#ifndef CTR_FWD_SLAB_S
#define CTR_FWD_SLAB_S 8
#endif
#ifndef CTR_FWD_SLAB_CAP
#define CTR_FWD_SLAB_CAP 6144
#endif
#ifndef CTR_FWD_SLAB_TU
#define CTR_FWD_SLAB_TU 16
#endif
// Tile order over the XCDs (workgroups reach the 8 XCDs round-robin by
// index): 0 plain (consecutive tiles on different XCDs); 1 each XCD takes one
// contiguous eighth of every view's tiles; 2 groups of G x G neighbouring
// tiles are dealt to the XCDs round-robin, a group's tiles consecutive on its
// XCD (neighbours' overlapping boxes meet in one L2, every XCD still covers
// the whole detector).  Measured (ms, the three inputs of the table above):
// 0: 347 / 22.4 / 1135; 1: 378 / 24.6 / 1065 (the XCDs holding the
// detector's edge rows run out of work early); 2 with G = 2: 344 / 22.2 /
// 1057, G = 3: 360 / 26.4 / 1054, G = 4: 347 / 24.5 / 1044.
#ifndef CTR_FWD_SLAB_XCD
#define CTR_FWD_SLAB_XCD 2
#endif
#ifndef CTR_FWD_SLAB_GROUP
#define CTR_FWD_SLAB_GROUP 2
#endif
constexpr int kSlabGroup = CTR_FWD_SLAB_GROUP;
// Consecutive views per block (1 or 2, chosen per call from the geometry):
// their rays through one detector tile cross nearly the same voxels, so one
// staged box (bounded by all their corner rays) serves them all.  A voxel r
// voxels from the rotation axis moves about r dtheta between views; the box
// grows by that much per extra view.  CTR_FWD_SLAB_VIEWS forces a count.
#ifndef CTR_FWD_SLAB_VIEWS
#define CTR_FWD_SLAB_VIEWS 0
#endif

// Blocks per view group of the slab grid (padding blocks included).
__host__ __device__ inline long long slab_blocks_per_view(int tu, int tv)
{
    const long long P = (long long)tu * tv;
    if (CTR_FWD_SLAB_XCD == 1) return 8 * ((P + 7) / 8);
    if (CTR_FWD_SLAB_XCD == 2) {
        const long long groups =
            (long long)((tu + kSlabGroup - 1) / kSlabGroup) * ((tv + kSlabGroup - 1) / kSlabGroup);
        return 8 * ((groups + 7) / 8) * kSlabGroup * kSlabGroup;
    }
    return P;
}
constexpr int kSlabS = CTR_FWD_SLAB_S;      // slices per slab
constexpr int kSlabCap = CTR_FWD_SLAB_CAP;  // floats of the staged box
constexpr int kSlabThreads = 256, kSlabWaves = kSlabThreads / 64;
constexpr int kSlabTU = CTR_FWD_SLAB_TU, kSlabTV = kSlabThreads / kSlabTU;
static_assert(kSlabTU * kSlabTV == kSlabThreads, "slab tile");

// trilinear0 out of line: the slab kernel's samples outside the staged box
// are rare, and inlined its eight bounds-checked taps would put a chain of
// exec-mask branches into every iteration of the sample loop.
__device__ __attribute__((noinline)) float trilinear0_call(const float *__restrict__ vol, int X,
                                                            int Y, int Z, float fi, float fj,
                                                            float fk)
{
    return trilinear0(vol, X, Y, Z, fi, fj, fk);
}

// A zero the box fill of volumes of 4 GiB and more reads for voxels outside
// the volume (device globals start zeroed; nothing writes it).
static __device__ float g_zero_word[1];

// q = x / d, r = x % d for 0 <= x < 2^16 and d > 0 (here x < 256): a float
// estimate of the quotient, off by at most one there even with the
// reciprocal one ulp off, and one correction step.
__device__ inline void small_divmod(int x, int d, int &q, int &r)
{
    q = (int)((float)x * __builtin_amdgcn_rcpf((float)d));
    r = x - q * d;
    if (r < 0) {
        --q;
        r += d;
    } else if (r >= d) {
        ++q;
        r -= d;
    }
}

__device__ inline int wave_min(int v)
{
    for (int off = 32; off > 0; off >>= 1) v = min(v, __shfl_xor(v, off));
    return v;
}
__device__ inline int wave_max(int v)
{
    for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off));
    return v;
}

// (x, y, z) -> (x Y + y) Z + z: the volume with z fastest.  Block (b, g)
// transposes 32 x 32 (z, x) tile b of the y planes g, g + gridDim.y, ...
// through LDS.
__global__ __launch_bounds__(256) void transpose_xz_kernel(const float *__restrict__ vol,
                                                           float *__restrict__ volT, int X, int Y,
                                                           int Z)
{
    __shared__ float t[32][33];
    const int tiles_x = (X + 31) / 32;
    const int x0 = (int)(blockIdx.x % tiles_x) * 32, z0 = (int)(blockIdx.x / tiles_x) * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
    for (int y = blockIdx.y; y < Y; y += gridDim.y) {
        for (int r = ty; r < 32; r += 8) {
            const int x = x0 + tx, z = z0 + r;
            if (x < X && z < Z) t[r][tx] = vol[((size_t)z * Y + y) * X + x];
        }
        __syncthreads();
        for (int r = ty; r < 32; r += 8) {
            const int z = z0 + tx, x = x0 + r;
            if (x < X && z < Z) volT[((size_t)x * Y + y) * Z + z] = t[tx][r];
        }
        __syncthreads();
    }
}

// The samples k, k + 1, ... of ray R whose cell starts in slab [lo, hi]
// along A, trilinear from the staged box ([A][y][C], C = z when AXX, else x,
// unit stride along C: the taps pair into two-dword LDS reads).  A sample
// whose cell leaves the box along y or C takes trilinear0 from memory.
template <bool AXX>
__device__ inline void slab_samples(const ForwardJob &J, const IndexRay &R, int &k, float &acc,
                                    const float *box, bool staged, int lo, int hi, int yl, int cl,
                                    int Ry, int Rc)
{
    const unsigned sb_a = (unsigned)(Ry * Rc), sj = (unsigned)Rc;
    const unsigned ry2 = staged ? (unsigned)(Ry - 2) : 0u, rc2 = (unsigned)(Rc - 2);
    // t = k + 1/2 as a float counter (exact: k < 2^23)
    const float tK = (float)R.K;
    float t = (float)k + 0.5f;
#pragma unroll 1
    while (t < tK) {
        const float fi = fmaf(t, R.s[0], R.b[0]), fj = fmaf(t, R.s[1], R.b[1]),
                    fk = fmaf(t, R.s[2], R.b[2]);
        const int ia = (int)floorf(AXX ? fi : fk);
        if (ia < lo || ia > hi) break;
        // the sample's cell in the staged box: along A it starts in the slab
        // (the test above); along y and C the box bounds come from FP32
        // t-ranges with a one-voxel margin, so a ray that is almost
        // perpendicular to A could still leave them -- such a sample takes
        // trilinear0 from memory (the same arithmetic, in order)
        const float i0f = floorf(fi), j0f = floorf(fj), k0f = floorf(fk);
        const int i0 = (int)i0f, j0 = (int)j0f, k0 = (int)k0f;
        const unsigned dy = (unsigned)(j0 - yl), dc = (unsigned)((AXX ? k0 : i0) - cl);
        const bool inbox = staged && dy <= ry2 && dc <= rc2;
        // trilinear0's arithmetic on the staged box (zero outside the
        // volume); the cell's 8 taps lie inside the S x Ry x Rc box.  Lanes
        // outside it read the box's first cell instead, and their value is
        // replaced by trilinear0's
        const float wi = fi - i0f, wj = fj - j0f, wk = fk - k0f;
        const unsigned ab_in = __umul24((unsigned)(ia - lo), sb_a) + __umul24(dy, sj) + dc;
        const float *p = box + (inbox ? ab_in : 0u);
        float t000, t100, t010, t110, t001, t101, t011, t111;  // t_ijk, offsets from the cell
        if constexpr (AXX) {  // C = z: unit stride in k, sb_a in i
            t000 = p[0], t001 = p[1], t010 = p[sj], t011 = p[sj + 1];
            t100 = p[sb_a], t101 = p[sb_a + 1], t110 = p[sb_a + sj], t111 = p[sb_a + sj + 1];
        } else {  // C = x: unit stride in i, sb_a in k
            t000 = p[0], t100 = p[1], t010 = p[sj], t110 = p[sj + 1];
            t001 = p[sb_a], t101 = p[sb_a + 1], t011 = p[sb_a + sj], t111 = p[sb_a + sj + 1];
        }
        const float c00 = fmaf(wi, t100 - t000, t000);
        const float c01 = fmaf(wi, t110 - t010, t010);
        const float c10 = fmaf(wi, t101 - t001, t001);
        const float c11 = fmaf(wi, t111 - t011, t011);
        const float c0 = fmaf(wj, c01 - c00, c00);
        const float c1 = fmaf(wj, c11 - c10, c10);
        float v = fmaf(wk, c1 - c0, c0);
        if (!inbox) v = trilinear0_call(J.vol, J.X, J.Y, J.Z, fi, fj, fk);
        acc += v;
        t += 1.0f;
    }
    k = (int)t;  // t = k + 1/2: truncation gives k
}

template <bool kBuf, int V>
__global__ __launch_bounds__(kSlabThreads) void forward_slab_kernel(ForwardJob J,
                                                                    const float *volT)
{
    __shared__ float box[kSlabCap];
    __shared__ int red[kSlabWaves][2];  // per-wave sample range along A
    __shared__ float4 corner[4 * V];    // corner rays: y, C = (Y0, dY, C0, dC) . (1, A)
    const int tu = (J.nu + kSlabTU - 1) / kSlabTU;
    const int tv = (J.nv + kSlabTV - 1) / kSlabTV;
    const int P = tu * tv;
    int ng, tile;  // view group (views ng V .. ng V + V - 1) and tile
    if (CTR_FWD_SLAB_XCD == 1) {
        // the grid holds 8 P8 blocks per view group: XCD x = index mod 8
        // takes tiles x P8 .. x P8 + P8 - 1 of each group in order; the
        // padding blocks return at once (no barrier reached)
        const int P8 = (P + 7) / 8;
        const int x = (int)(blockIdx.x & 7), kk = (int)(blockIdx.x >> 3);
        ng = kk / P8;
        tile = x * P8 + kk % P8;
        if (tile >= P) return;
    } else if (CTR_FWD_SLAB_XCD == 2) {
        // block b: XCD x = b mod 8, its j-th block j = b / 8 = (view group,
        // group slot, tile in group); tile group g = slot * 8 + x of the
        // G x G groups (row-major over the tile grid); padding blocks return
        constexpr int G2 = kSlabGroup * kSlabGroup;
        const int gu = (tu + kSlabGroup - 1) / kSlabGroup;
        const int groups = gu * ((tv + kSlabGroup - 1) / kSlabGroup);
        const int per_x = (groups + 7) / 8;  // group slots per XCD and view group
        const int x = (int)(blockIdx.x & 7), j = (int)(blockIdx.x >> 3);
        ng = j / (per_x * G2);
        const int rem = j - ng * per_x * G2;
        const int g = (rem / G2) * 8 + x, t = rem % G2;
        const int u = (g % gu) * kSlabGroup + t % kSlabGroup;
        const int v = (g / gu) * kSlabGroup + t / kSlabGroup;
        if (g >= groups || u >= tu || v >= tv) return;
        tile = v * tu + u;
    } else {
        ng = blockIdx.x / P;
        tile = blockIdx.x % P;
    }
    const int n0 = ng * V;
    const int iu = (tile % tu) * kSlabTU + (int)(threadIdx.x % kSlabTU);
    const int iv = (tile / tu) * kSlabTV + (int)(threadIdx.x / kSlabTU);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const bool pix = iu < J.nu && iv < J.nv;
    // this pixel's ray in each of the block's views
    IndexRay R[V];
#pragma unroll
    for (int v = 0; v < V; ++v) {
        R[v].K = 0;
        for (int a = 0; a < 3; a++) R[v].b[a] = R[v].s[a] = 0.0f;
        if (pix && n0 + v < J.n_views) R[v] = index_ray(J, n0 + v, iu, iv);
    }

    // the main horizontal axis of the block's first view and its central
    // ray's direction on it (the block's other views, a fraction of a degree
    // further on, march along the same axis; a ray running against it takes
    // its samples after the slabs)
    const double th = 2.0 * M_PI * J.rot * (double)n0 / (double)J.n_views + J.offset;
    double sn, cs;
    sincos(th, &sn, &cs);
    // (without the copy every view marches along z: slower for views near
    // the x axis, the same sums)
    const bool ax_x = volT != nullptr && fabs(cs) >= fabs(sn);
    const bool up = ax_x ? cs < 0.0 : sn < 0.0;
    const int dimA = ax_x ? J.X : J.Z, dimC = ax_x ? J.Z : J.X;
    bool along[V];
#pragma unroll
    for (int v = 0; v < V; ++v) {
        const float sA = ax_x ? R[v].s[0] : R[v].s[2];
        along[v] = up ? sA > 0.0f : sA < 0.0f;  // else: all samples after the slabs
    }

    // once per block: the cells along A of the along-rays' first and last
    // samples, and the tile's four corner rays in each view.  Every ray of
    // the tile lies in the convex hull of its view's corner rays, so between
    // two planes of constant A its points lie in the bounding box of those
    // corner rays' points there: the staged box of a slab (a sample outside
    // it still takes trilinear0, so the bounds only size the box)
    {
        int amin = 1 << 30, amax = -(1 << 30);
#pragma unroll
        for (int v = 0; v < V; ++v) {
            if (along[v] && R[v].K > 0) {
                const float bA = ax_x ? R[v].b[0] : R[v].b[2], sA = ax_x ? R[v].s[0] : R[v].s[2];
                const float f0 = fmaf(0.5f, sA, bA), f1 = fmaf((float)R[v].K - 0.5f, sA, bA);
                amin = min(amin, (int)floorf(fminf(f0, f1)));
                amax = max(amax, (int)floorf(fmaxf(f0, f1)));
            }
        }
        amin = wave_min(amin);
        amax = wave_max(amax);
        if (lane == 0) {
            red[wave][0] = amin;
            red[wave][1] = amax;
        }
        if (threadIdx.x < 4 * V) {
            const int q = threadIdx.x & 3, nv_ = min(n0 + (int)threadIdx.x / 4, J.n_views - 1);
            const int u0 = (tile % tu) * kSlabTU, v0 = (tile / tu) * kSlabTV;
            const int cu = min(u0 + (q & 1) * (kSlabTU - 1), J.nu - 1);
            const int cv = min(v0 + (q >> 1) * (kSlabTV - 1), J.nv - 1);
            float L[6];
            index_line(J, nv_, cu, cv, L);
            // y and C of the ray as functions of A: y = Y0 + A dY, C = C0 + A dC
            const float cA = ax_x ? L[0] : L[2], dA = ax_x ? L[3] : L[5];
            const float dY = L[4] / dA, dC = (ax_x ? L[5] : L[3]) / dA;
            corner[threadIdx.x] = make_float4(fmaf(-cA, dY, L[1]), dY,
                                              fmaf(-cA, dC, ax_x ? L[2] : L[0]), dC);
        }
    }
    __syncthreads();
    int amin = red[0][0], amax = red[0][1];
#pragma unroll
    for (int w = 1; w < kSlabWaves; ++w) {
        amin = min(amin, red[w][0]);
        amax = max(amax, red[w][1]);
    }

    // box rows come from (a Y + y) dimC + c of the volume (A = z) or of its
    // z-fastest copy (A = x): through a buffer resource below 4 GiB, 64-bit
    // addresses above
    const float *src = ax_x ? volT : J.vol;
    const unsigned vol_bytes = kBuf ? (unsigned)J.X * (unsigned)J.Y * (unsigned)J.Z * 4u : 0u;
    const __amdgpu_buffer_rsrc_t vrs = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<float *>(src), (short)0, (int)vol_bytes, 0x00020000);

    // cells start at index -2 .. dimA along A (samples lie in the box grown
    // by one voxel).  The slabs cover the cells of the block's samples,
    // amin .. amax, in the march direction: a slab of S slices holds cell
    // starts lo .. lo + S - 2 and slices lo .. lo + S - 1, S = kSlabS unless
    // its box would exceed the LDS region (wide detector pixels, views far
    // apart): then the deepest slab that fits, down to S = 2.
    float acc[V];
    int k[V];
#pragma unroll
    for (int v = 0; v < V; ++v) {
        acc[v] = 0.0f;
        k[v] = 0;
    }
    const int cfirst = max(amin, -2), clast = min(amax, dimA);
    int cur = up ? cfirst : clast;  // next cell start to cover
#pragma unroll 1
    while (up ? cur <= clast : cur >= cfirst) {
        __syncthreads();  // every ray is done with the previous slab's box
        int S = kSlabS, lo = 0, hi = 0, yl = 0, yh = -1, cl = 0, ch = -1;
#pragma unroll 1
        for (int attempt = 0; attempt < 2; ++attempt) {
            lo = up ? cur : cur - (S - 2);
            hi = lo + S - 2;
            // the corner rays' y and C on the planes A = lo and A = hi + 1
            // (the slab's cells start in [lo, hi + 1)), widened by one voxel
            // for rounding and clamped to the cells samples can have (-2 ..
            // n); the lines are read from LDS, not kept in registers
            float ylo = 3.0e38f, yhi = -3.0e38f, clo = 3.0e38f, chi = -3.0e38f;
#pragma unroll
            for (int q = 0; q < 4 * V; ++q) {
                const float4 L = corner[q];
#pragma unroll
                for (int f = 0; f < 2; ++f) {
                    const float pa = (float)(f ? hi + 1 : lo);
                    const float y = fmaf(pa, L.y, L.x), c = fmaf(pa, L.w, L.z);
                    ylo = fminf(ylo, y);
                    yhi = fmaxf(yhi, y);
                    clo = fminf(clo, c);
                    chi = fmaxf(chi, c);
                }
            }
            // a corner ray (almost) parallel to the planes gives no finite bound
            if (!(ylo <= yhi && clo <= chi)) {
                ylo = -4.0f;
                yhi = (float)J.Y + 4.0f;
                clo = -4.0f;
                chi = (float)dimC + 4.0f;
            }
            const float ymax = (float)J.Y + 4.0f, cmax = (float)dimC + 4.0f;
            yl = max((int)floorf(fminf(fmaxf(ylo, -4.0f), ymax)) - 1, -2);
            yh = min((int)floorf(fminf(fmaxf(yhi, -4.0f), ymax)) + 2, J.Y + 1);
            cl = max((int)floorf(fminf(fmaxf(clo, -4.0f), cmax)) - 1, -2);
            ch = min((int)floorf(fminf(fmaxf(chi, -4.0f), cmax)) + 2, dimC + 1);
            const long long face = (long long)max(yh - yl + 1, 1) * max(ch - cl + 1, 1);
            // (lo, hi and the bounds stay those of the last S tried: a box
            // that still does not fit reads the volume directly)
            if ((long long)S * face <= kSlabCap || S == 2 || attempt == 1) break;
            S = (int)max(2LL, min((long long)S - 1, kSlabCap / face));
        }
        cur = up ? hi + 1 : lo - 1;
        if (yh < yl || ch < cl) continue;  // block-uniform
        const int Ry = yh - yl + 1, Rc = ch - cl + 1;
        const bool staged = (long long)S * Ry * Rc <= kSlabCap;
        if (staged) {
            // element e = (a' Ry + y') Rc + c' of the box, the thread's
            // (a', y', c') stepped by kSlabThreads elements at a time
            const int total = S * Ry * Rc;
            int q, qc, qa, qy, r0, c, a, y;
            small_divmod(kSlabThreads, Rc, q, qc);
            small_divmod(q, Ry, qa, qy);
            small_divmod((int)threadIdx.x, Rc, r0, c);
            small_divmod(r0, Ry, a, y);
            for (int e = threadIdx.x; e < total; e += kSlabThreads) {
                const int ga = lo + a, gy = yl + y, gc = cl + c;
                const bool in = (unsigned)ga < (unsigned)dimA && (unsigned)gy < (unsigned)J.Y &&
                                (unsigned)gc < (unsigned)dimC;
                // LDS DMA: the wave's lanes stage 64 consecutive box elements
                // from e - lane on; a voxel outside the volume reads zero
                // through an offset past the buffer's end (below 4 GiB) or
                // from g_zero_word (64-bit addresses)
                const int ebase = __builtin_amdgcn_readfirstlane(e - lane);
                auto *dst = (__attribute__((address_space(3))) void *)(box + ebase);
                if constexpr (kBuf) {
                    const unsigned voff =
                        in ? (unsigned)((ga * J.Y + gy) * dimC + gc) * 4u : 0xfffffffcu;
                    __builtin_amdgcn_raw_ptr_buffer_load_lds(vrs, dst, 4, voff, 0, 0, 0);
                } else {
                    const float *gp =
                        in ? src + (((size_t)ga * J.Y + gy) * dimC + gc) : g_zero_word;
                    __builtin_amdgcn_global_load_lds(
                        (const __attribute__((address_space(1))) void *)gp, dst, 4, 0, 0);
                }
                c += qc;
                int carry = 0;
                if (c >= Rc) {
                    c -= Rc;
                    carry = 1;
                }
                y += qy + carry;
                a += qa;
                if (y >= Ry) {
                    y -= Ry;
                    ++a;
                }
            }
            asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        }
        __syncthreads();
#pragma unroll
        for (int v = 0; v < V; ++v) {
            if (!along[v]) continue;
            if (ax_x)
                slab_samples<true>(J, R[v], k[v], acc[v], box, staged, lo, hi, yl, cl, Ry, Rc);
            else
                slab_samples<false>(J, R[v], k[v], acc[v], box, staged, lo, hi, yl, cl, Ry, Rc);
        }
    }
    if (!pix) return;
#pragma unroll
    for (int v = 0; v < V; ++v) {
        const int n = n0 + v;
        if (n >= J.n_views) break;
        for (; k[v] < R[v].K; ++k[v]) {
            const float t = (float)k[v] + 0.5f;
            acc[v] += trilinear0(J.vol, J.X, J.Y, J.Z, fmaf(t, R[v].s[0], R[v].b[0]),
                                 fmaf(t, R[v].s[1], R[v].b[1]), fmaf(t, R[v].s[2], R[v].b[2]));
        }
        if (R[v].K > 0) acc[v] *= (float)J.h;
        const size_t idx = J.layout == CTR_LAYOUT_PROJECTIONS
                               ? ((size_t)n * J.nv + iv) * J.nu + iu
                               : ((size_t)iv * J.n_views + n) * J.nu + iu;
        J.out[idx] = acc[v];
    }
}

// ---------------------------------------------------------------------------
// Fan -> parallel rebinning (FBP2D::parallelize_fan_beam_sinogram)
// ---------------------------------------------------------------------------
__global__ void rebin_kernel(const float *fan, int nu_fan, int n_views, double sod, double sdd,
                             double delta, double rot, float *par, int nu_par, double pitch_par)
{
    long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)nu_par * n_views) return;
    int is = (int)(idx % nu_par);
    int n = (int)(idx / nu_par);
    double dth = 2.0 * M_PI * rot / (double)n_views;
    double theta = dth * (double)n;
    double s = (is + 0.5 - 0.5 * nu_par) * pitch_par;
    float out = 0.0f;
    if (fabs(s) < sod) {
        double gam = asin(s / sod);
        double u = sdd * tan(gam);
        double beta = theta + gam;
        double fb = beta / dth;
        double fu = u / delta + 0.5 * nu_fan - 0.5;
        bool periodic = fabs(rot - 1.0) < 1e-12;
        if (fu > -1.0 && fu < (double)nu_fan) {
            double fb0 = floor(fb), fu0 = floor(fu);
            double wb = fb - fb0, wu = fu - fu0;
            double accum = 0.0;
            for (int b = 0; b < 2; b++) {
                long long vb = (long long)fb0 + b;
                if (periodic) {
                    vb %= n_views;
                    if (vb < 0) vb += n_views;
                } else if (vb < 0 || vb >= n_views) {
                    continue;
                }
                double wbv = b ? wb : 1.0 - wb;
                for (int a = 0; a < 2; a++) {
                    int cu = (int)fu0 + a;
                    if (cu < 0 || cu >= nu_fan) continue;
                    double wa = a ? wu : 1.0 - wu;
                    accum += wbv * wa * (double)fan[(size_t)vb * nu_fan + cu];
                }
            }
            out = (float)accum;
        }
    }
    par[idx] = out;
}

// ---------------------------------------------------------------------------
// Elementwise passes
// ---------------------------------------------------------------------------
__global__ void fdk_weight_kernel(float *proj, int layout, int nu, int nv, int n_views,
                                  double sdd, double delta)
{
    long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long total = (long long)nu * nv * n_views;
    if (idx >= total) return;
    int iu = (int)(idx % nu);
    long long r = idx / nu;
    int iv = layout == CTR_LAYOUT_PROJECTIONS ? (int)(r % nv) : (int)(r / n_views);
    double u = (iu + 0.5 - 0.5 * nu) * delta, v = (iv + 0.5 - 0.5 * nv) * delta;
    proj[idx] = (float)((double)proj[idx] * (sdd / sqrt(sdd * sdd + u * u + v * v)));
}

__global__ void log_normalize_kernel(float *p, long long count, float base, float floor_value)
{
    long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= count) return;
    p[idx] = -logf(fmaxf(p[idx], floor_value) / base);
}

__global__ void clip_kernel(float *p, long long count)
{
    long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx < count) p[idx] = fmaxf(p[idx], 0.0f);
}

__global__ void fill_kernel(float *p, long long count, float value)
{
    long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx < count) p[idx] = value;
}

// Ellipsoid record prepared on the host: centre (3), body-frame rows of
// R^T scaled by 1/semi-axis (9), density (1).
struct EllipsoidRec {
    double c[3];
    double m[9];
    double density;
};

__global__ void render_kernel(const EllipsoidRec *E, int count, float *vol, int X, int Y, int Z,
                              double d)
{
    long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)X * Y * Z) return;
    int i = (int)(idx % X);
    int j = (int)((idx / X) % Y);
    int k = (int)(idx / ((long long)X * Y));
    double x = (i + 0.5 - 0.5 * X) * d, y = (j + 0.5 - 0.5 * Y) * d, z = (k + 0.5 - 0.5 * Z) * d;
    float acc = vol[idx];
    for (int e = 0; e < count; e++) {
        const EllipsoidRec &R = E[e];
        double p0 = x - R.c[0], p1 = y - R.c[1], p2 = z - R.c[2];
        double q = 0.0;
        for (int a = 0; a < 3; a++) {
            double b = R.m[3 * a] * p0 + R.m[3 * a + 1] * p1 + R.m[3 * a + 2] * p2;
            q += b * b;
        }
        if (q <= 1.0) acc += (float)R.density;
    }
    vol[idx] = acc;
}

// Counter-based generator: splitmix64 stream keyed by (seed, index).
struct SplitMix {
    unsigned long long s;
    __device__ inline unsigned long long next()
    {
        unsigned long long z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    __device__ inline double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }
};

// Poisson sample: multiplication method below mean 10, otherwise Hormann's
// transformed rejection with squeeze (PTRS, 1993).  Loops are bounded.
__device__ inline double poisson(SplitMix &g, double mu)
{
    if (mu <= 0.0) return 0.0;
    if (mu < 10.0) {
        double L = exp(-mu), p = 1.0;
        int k = 0;
        for (; k < 1000; ++k) {
            p *= g.uniform();
            if (p <= L) break;
        }
        return (double)k;
    }
    double smu = sqrt(mu), b = 0.931 + 2.53 * smu, a = -0.059 + 0.02483 * b;
    double inv_alpha = 1.1239 + 1.1328 / (b - 3.4), vr = 0.9277 - 3.6224 / (b - 2.0);
    double lmu = log(mu);
    for (int it = 0; it < 256; ++it) {
        double U = g.uniform() - 0.5, V = g.uniform();
        double us = 0.5 - fabs(U);
        double k = floor((2.0 * a / us + b) * U + mu + 0.43);
        if (us >= 0.07 && V <= vr) return k;
        if (k < 0.0 || (us < 0.013 && V > us)) continue;
        if (log(V) + log(inv_alpha) - log(a / (us * us) + b) <= -mu + k * lmu - lgamma(k + 1.0))
            return k;
    }
    return floor(mu + 0.5);
}

__global__ void poisson_log_kernel(float *p, long long count, double i0, unsigned long long seed)
{
    long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= count) return;
    SplitMix g{seed ^ ((unsigned long long)idx * 0xD1B54A32D192ED03ull)};
    g.next();
    double mu = i0 * exp(-(double)p[idx]);
    double k = poisson(g, mu);
    p[idx] = (float)(-log(fmax(k, 1.0) / i0));
}

// Host side of both ellipsoid generators: a record of 10 doubles (centre,
// semi-axes, angles alpha beta gamma, density) -> centre, the rows of R^T
// divided by the semi-axes (world -> unit-sphere body frame), with
// R = Rz(gamma) Ry(beta) Rx(alpha), and the density.
static int prepare_ellipsoid(const double *P, int e, EllipsoidRec *Q)
{
    CTR_REQUIRE(P[3] > 0 && P[4] > 0 && P[5] > 0, CTR_ERR_INVALID_ARGUMENT,
                "ellipsoid %d has a non-positive semi-axis", e);
    double ca = cos(P[6]), sa = sin(P[6]), cb = cos(P[7]), sb = sin(P[7]);
    double cg = cos(P[8]), sg = sin(P[8]);
    double Rx[3][3] = {{1, 0, 0}, {0, ca, -sa}, {0, sa, ca}};
    double Ry[3][3] = {{cb, 0, sb}, {0, 1, 0}, {-sb, 0, cb}};
    double Rz[3][3] = {{cg, -sg, 0}, {sg, cg, 0}, {0, 0, 1}};
    double T[3][3], R[3][3];
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) {
            T[r][c] = 0;
            for (int m = 0; m < 3; m++) T[r][c] += Ry[r][m] * Rx[m][c];
        }
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) {
            R[r][c] = 0;
            for (int m = 0; m < 3; m++) R[r][c] += Rz[r][m] * T[m][c];
        }
    for (int a = 0; a < 3; a++) {
        Q->c[a] = P[a];
        for (int b = 0; b < 3; b++) Q->m[3 * a + b] = R[b][a] / P[3 + a];
    }
    Q->density = P[9];
    return CTR_OK;
}

// ---------------------------------------------------------------------------
// Analytic projections of ellipsoids: the exact line integral
// sum_e density_e * chord_e along each detector ray, the rays of
// ctr_forward_project (source SOD (cos t, 0, sin t), flat detector at the
// detector distance beyond the axis; parallel rays along -(cos t, 0, sin t)
// through u (-sin t, 0, cos t) + v y).  One workgroup per 16 x 16 pixel tile
// of one view: one thread computes the view's sine and cosine and the
// tile's ray cone, one thread per ellipsoid culls its bounding sphere
// against that cone (FP64, conservative), and every ray then takes only the
// candidates of its tile.  The chord is FP64: the ray is moved to its point
// nearest the centre, taken to the unit-sphere body frame, and the chord is
// 2 sqrt(B^2 - A C) / A of |r + s q|^2 = 1.  The records travel in the
// kernel argument (no device table, no host wait): kEllGroup per launch.
// ---------------------------------------------------------------------------
constexpr int kEllGroup = 32;  // one bit each in the tile's candidate mask
constexpr int kEllTile = 16;
struct EllipsoidGroup {
    double c[kEllGroup][3];
    double m[kEllGroup][9];
    double density[kEllGroup];
    double radius[kEllGroup];  // bounding sphere: the largest semi-axis
    int count;
};
struct RayGeometry {
    int geometry, layout, nu, nv, n_views, accumulate;
    double sod, odd, delta, rot, offset;
    float *out;
};
static_assert(sizeof(EllipsoidGroup) + sizeof(RayGeometry) < 4000, "kernel argument size");

__device__ inline double dot3(const double *a, const double *b)
{
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

// Unit direction from the source to detector point (u, v) of a cone view.
__device__ inline void cone_ray(const RayGeometry &G, double st, double ct, double u, double v,
                                double *d)
{
    d[0] = -G.odd * ct - u * st - G.sod * ct;
    d[1] = v;
    d[2] = -G.odd * st + u * ct - G.sod * st;
    const double inv = 1.0 / sqrt(dot3(d, d));
    d[0] *= inv;
    d[1] *= inv;
    d[2] *= inv;
}

__global__ __launch_bounds__(256) void project_ellipsoids_kernel(EllipsoidGroup E, RayGeometry G)
{
    // view trig; the tile's ray cone (cone beam: axis, cos and sin of its
    // half-angle) or pixel-centre rectangle (parallel beam)
    __shared__ double tile[8];
    __shared__ unsigned cand;
    const int tu = (G.nu + kEllTile - 1) / kEllTile, tv = (G.nv + kEllTile - 1) / kEllTile;
    const int n = (int)(blockIdx.x / (unsigned)(tu * tv));
    const int t = (int)(blockIdx.x % (unsigned)(tu * tv));
    const int u0 = (t % tu) * kEllTile, v0 = (t / tu) * kEllTile;
    const int u1 = min(u0 + kEllTile, G.nu) - 1, v1 = min(v0 + kEllTile, G.nv) - 1;
    const bool cone = G.geometry == CTR_GEOMETRY_CONE;
    if (threadIdx.x == 0) {
        // FBP3D.cpp:223-227,283-284: theta_n = 2 pi rot n / N + offset
        const double th = 2.0 * M_PI * G.rot * (double)n / (double)G.n_views + G.offset;
        double st, ct;
        sincos(th, &st, &ct);
        tile[0] = st;
        tile[1] = ct;
        const double ua = (u0 + 0.5 - 0.5 * G.nu) * G.delta, ub = (u1 + 0.5 - 0.5 * G.nu) * G.delta;
        const double va = (v0 + 0.5 - 0.5 * G.nv) * G.delta, vb = (v1 + 0.5 - 0.5 * G.nv) * G.delta;
        if (cone) {
            double axis[3], d[3];
            cone_ray(G, st, ct, 0.5 * (ua + ub), 0.5 * (va + vb), axis);
            double cmin = 1.0;
            for (int k = 0; k < 4; k++) {
                cone_ray(G, st, ct, k & 1 ? ub : ua, k & 2 ? vb : va, d);
                cmin = fmin(cmin, dot3(axis, d));
            }
            tile[2] = axis[0];
            tile[3] = axis[1];
            tile[4] = axis[2];
            tile[5] = cmin;                                   // cos(beta)
            tile[6] = sqrt(fmax(0.0, 1.0 - cmin * cmin));     // sin(beta)
        } else {
            tile[2] = ua;
            tile[3] = ub;
            tile[4] = va;
            tile[5] = vb;
        }
        cand = 0u;
    }
    __syncthreads();
    const double st = tile[0], ct = tile[1];
    if ((int)threadIdx.x < E.count) {
        const int e = threadIdx.x;
        const double R = E.radius[e] * (1.0 + 1e-9) + 1e-9;
        bool hit;
        if (cone) {
            // a line of the tile's ray cone can meet the sphere only if the
            // angle between the cone axis and the centre direction is at most
            // alpha + beta (alpha the sphere's angular radius)
            const double w[3] = {E.c[e][0] - G.sod * ct, E.c[e][1], E.c[e][2] - G.sod * st};
            const double L = sqrt(dot3(w, w));
            if (L <= R) {
                hit = true;
            } else {
                const double axis[3] = {tile[2], tile[3], tile[4]};
                const double sa = R / L, ca = sqrt(1.0 - sa * sa);
                const double cb = tile[5], sb = tile[6];
                const double cg = dot3(axis, w) / L;
                // cos(alpha + beta), or -1 once alpha + beta reaches pi
                const double cab = (ca * cb - sa * sb < -1.0 || (ca < 0.0 && cb < 0.0))
                                       ? -1.0
                                       : ca * cb - sa * sb;
                // a line integral: the cone about -axis counts too
                hit = fabs(cg) >= cab - 1e-12;
            }
        } else {
            // nearest point of the tile's pixel-centre rectangle to the
            // centre's detector coordinates
            const double uc = -E.c[e][0] * st + E.c[e][2] * ct, vc = E.c[e][1];
            const double du = fmax(fmax(tile[2] - uc, uc - tile[3]), 0.0);
            const double dv = fmax(fmax(tile[4] - vc, vc - tile[5]), 0.0);
            hit = du * du + dv * dv <= R * R;
        }
        if (hit) atomicOr(&cand, 1u << e);
    }
    __syncthreads();
    unsigned mask = cand;
    const int iu = u0 + (int)(threadIdx.x % kEllTile), iv = v0 + (int)(threadIdx.x / kEllTile);
    if (iu >= G.nu || iv >= G.nv) return;
    const double u = (iu + 0.5 - 0.5 * G.nu) * G.delta, v = (iv + 0.5 - 0.5 * G.nv) * G.delta;
    double o[3], d[3];
    if (cone) {
        o[0] = G.sod * ct;
        o[1] = 0.0;
        o[2] = G.sod * st;
        cone_ray(G, st, ct, u, v, d);
    } else {
        o[0] = -u * st;
        o[1] = v;
        o[2] = u * ct;
        d[0] = -ct;
        d[1] = 0.0;
        d[2] = -st;
    }
    double acc = 0.0;
    while (mask) {
        const int e = __ffs(mask) - 1;
        mask &= mask - 1;
        // p = point of the ray nearest the centre, minus the centre
        const double w0 = E.c[e][0] - o[0], w1 = E.c[e][1] - o[1], w2 = E.c[e][2] - o[2];
        const double t0 = w0 * d[0] + w1 * d[1] + w2 * d[2];
        const double p0 = t0 * d[0] - w0, p1 = t0 * d[1] - w1, p2 = t0 * d[2] - w2;
        const double *M = E.m[e];
        double A = 0.0, B = 0.0, C = -1.0;
        for (int a = 0; a < 3; a++) {
            const double ra = M[3 * a] * p0 + M[3 * a + 1] * p1 + M[3 * a + 2] * p2;
            const double qa = M[3 * a] * d[0] + M[3 * a + 1] * d[1] + M[3 * a + 2] * d[2];
            A += qa * qa;
            B += ra * qa;
            C += ra * ra;
        }
        const double disc = B * B - A * C;
        if (disc > 0.0) acc += E.density[e] * 2.0 * sqrt(disc) / A;
    }
    const size_t idx = G.layout == CTR_LAYOUT_PROJECTIONS
                           ? ((size_t)n * G.nv + iv) * G.nu + iu
                           : ((size_t)iv * G.n_views + n) * G.nu + iu;
    float *dst = G.out + idx;
    *dst = G.accumulate ? *dst + (float)acc : (float)acc;
}

}  // namespace ctr

using namespace ctr;

extern "C" {

int ctr_forward_project(void *stream, ctr_workspace *ws, int geometry, const float *volume,
                        int X, int Y, int Z, double source_distance, double detector_distance,
                        double detector_width, double volume_width, double rotation_count,
                        double rotation_offset, float *out, int layout, int nu, int nv,
                        int n_views, double step)
{
    clear_error();
    CTR_REQUIRE(volume && out, CTR_ERR_NULL_POINTER,
                "[CTAnalyzer Error] ctr_forward_project: volume or output is NULL");
    CTR_REQUIRE(layout == CTR_LAYOUT_PROJECTIONS || layout == CTR_LAYOUT_SINOGRAMS,
                CTR_ERR_INVALID_ARGUMENT, "unknown layout %d", layout);
    CTR_REQUIRE(step > 0, CTR_ERR_INVALID_ARGUMENT, "step must be positive");
    int rc = check_geometry(geometry, source_distance, detector_distance, detector_width,
                            volume_width, X, Y, Z, nu, nv, n_views);
    if (rc) return rc;
    ctr_workspace *w = nullptr;
    if ((rc = workspace_resolve(ws, &w))) return rc;
    ForwardJob J;
    J.geometry = geometry;
    J.vol = volume;
    J.X = X;
    J.Y = Y;
    J.Z = Z;
    J.d = volume_width / X;
    J.sod = source_distance;
    J.odd = detector_distance;
    J.delta = detector_width / nu;
    J.rot = rotation_count;
    J.offset = rotation_offset;
    J.h = step * J.d;
    J.out = out;
    J.layout = layout;
    J.nu = nu;
    J.nv = nv;
    J.n_views = n_views;
    const long long blocks = (long long)((nu + CTR_FWD_BU - 1) / CTR_FWD_BU) *
                             ((nv + 256 / CTR_FWD_BU - 1) / (256 / CTR_FWD_BU)) * n_views;
    CTR_REQUIRE(blocks < (1LL << 31), CTR_ERR_INVALID_ARGUMENT, "forward projection too large");
    const dim3 grid((unsigned)blocks);
    // interior cells through a buffer resource: byte offsets below 4 GiB and
    // the 24-bit products of the cell offset (the workspace option
    // CTR_OPTION_FORWARD_ADDRESSING = CTR_KERNEL_POINTER forces the 64-bit
    // path: diagnostics and tests)
    const unsigned long long bytes = 4ull * (unsigned long long)X * Y * Z;
    const bool buf = bytes <= 0xfffffff0ull && (unsigned long long)X * Y < (1ull << 24) &&
                     X < (1 << 24) && Y < (1 << 24) && Z < (1 << 24) &&
                     workspace_option(w, CTR_OPTION_FORWARD_ADDRESSING) != CTR_KERNEL_POINTER;
    // the slab-staged kernel by default, its boxes filled through the same
    // buffer resource or 64-bit addresses (CTR_OPTION_FORWARD_KERNEL =
    // CTR_KERNEL_RAY forces the ray kernel: diagnostics and tests)
    const bool slab = workspace_option(w, CTR_OPTION_FORWARD_KERNEL) != CTR_KERNEL_RAY;
    // views per block: the shift r dtheta of the volume's corner voxels
    // between consecutive views, r = half the volume's (x, z) diagonal in
    // voxels.  Measured (ms; 1 / 2 / 4 views per block; the shift in
    // brackets): 512^3 x 720 (3.2) 377 / 266 / 273, 256^3 x 360 (3.2) 24.6 /
    // 17.6 / 18.0, 1024^3 x 1200 (3.8) 3088 / 2073 / 1980, 1440^3 x 720 (8.9)
    // 5619 / 4191 / 4446, x 360 (17.8) 2812 / 2262 / 2764, x 96 (66) 755 /
    // 1027 / 2099; four views (127 VGPRs) are not kept.
    const double shift = 0.5 * sqrt((double)X * X + (double)Z * Z) * 2.0 * M_PI *
                         fabs(rotation_count) / (double)n_views;
    int views = shift <= 30.0 ? 2 : 1;
    if (CTR_FWD_SLAB_VIEWS == 1 || CTR_FWD_SLAB_VIEWS == 2) views = CTR_FWD_SLAB_VIEWS;
    const long long slab_blocks =
        slab_blocks_per_view((nu + kSlabTU - 1) / kSlabTU, (nv + kSlabTV - 1) / kSlabTV) *
        ((n_views + views - 1) / views);
    CTR_REQUIRE(slab_blocks < (1LL << 31), CTR_ERR_INVALID_ARGUMENT,
                "forward projection too large");
    const dim3 slab_grid((unsigned)slab_blocks);
    const float *volT = nullptr;
    if (slab) {
        // the z-fastest copy the slab kernel reads for views marching along
        // x (|cos theta| >= |sin theta| on the device; the margin covers host
        // and device rounding), rebuilt on every call
        bool any_x = false;
        for (int n = 0; n < n_views && !any_x; ++n) {
            const double th = 2.0 * M_PI * rotation_count * (double)n / (double)n_views +
                              rotation_offset;
            any_x = fabs(cos(th)) >= fabs(sin(th)) - 1e-6;
        }
        if (any_x) {
            void *t = nullptr;
            if ((rc = workspace_get(w, SLOT_VOLUME_T, bytes, &t))) return rc;
            const dim3 tg((unsigned)(((X + 31) / 32) * ((Z + 31) / 32)),
                          (unsigned)(Y < 65535 ? Y : 65535));
            transpose_xz_kernel<<<tg, 256, 0, as_stream(stream)>>>(volume, (float *)t, X, Y, Z);
            CTR_CHECK_LAUNCH();
            volT = (const float *)t;
        }
    }
#define CTR_SLAB_LAUNCH(B, V)                                                              \
    forward_slab_kernel<B, V><<<slab_grid, kSlabThreads, 0, as_stream(stream)>>>(J, volT)
    if (slab && buf) {
        if (views == 2) CTR_SLAB_LAUNCH(true, 2);
        else CTR_SLAB_LAUNCH(true, 1);
    } else if (slab) {
        if (views == 2) CTR_SLAB_LAUNCH(false, 2);
        else CTR_SLAB_LAUNCH(false, 1);
    } else if (buf) {
        forward_kernel<true><<<grid, 256, 0, as_stream(stream)>>>(J);
    } else {
        forward_kernel<false><<<grid, 256, 0, as_stream(stream)>>>(J);
    }
#undef CTR_SLAB_LAUNCH
    CTR_CHECK_LAUNCH();
    return CTR_OK;
}

int ctr_rebin_fan_to_parallel(void *stream, const float *fan, int nu_fan, int n_views,
                              double source_distance, double detector_distance,
                              double detector_width, double rotation_count, float *parallel,
                              int nu_par, double parallel_width)
{
    clear_error();
    CTR_REQUIRE(fan && parallel, CTR_ERR_NULL_POINTER, "ctr_rebin_fan_to_parallel: NULL buffer");
    CTR_REQUIRE(nu_fan > 0 && n_views > 0 && nu_par > 0, CTR_ERR_INVALID_ARGUMENT,
                "ctr_rebin_fan_to_parallel: sizes must be positive");
    CTR_REQUIRE(source_distance > 0 && detector_distance >= 0 && detector_width > 0 &&
                    parallel_width > 0 && rotation_count != 0.0,
                CTR_ERR_INVALID_ARGUMENT, "ctr_rebin_fan_to_parallel: bad geometry");
    long long total = (long long)nu_par * n_views;
    rebin_kernel<<<grid1d(total, 256), 256, 0, as_stream(stream)>>>(
        fan, nu_fan, n_views, source_distance, source_distance + detector_distance,
        detector_width / nu_fan, rotation_count, parallel, nu_par, parallel_width / nu_par);
    CTR_CHECK_LAUNCH();
    return CTR_OK;
}

int ctr_fdk_weight(void *stream, float *proj, int layout, int nu, int nv, int n_views,
                   double source_distance, double detector_distance, double detector_width)
{
    clear_error();
    CTR_REQUIRE(proj, CTR_ERR_NULL_POINTER, "ctr_fdk_weight: projections are NULL");
    CTR_REQUIRE(nu > 0 && nv > 0 && n_views >= 0, CTR_ERR_INVALID_ARGUMENT,
                "ctr_fdk_weight: bad sizes");
    CTR_REQUIRE(layout == CTR_LAYOUT_PROJECTIONS || layout == CTR_LAYOUT_SINOGRAMS,
                CTR_ERR_INVALID_ARGUMENT, "unknown layout %d", layout);
    CTR_REQUIRE(source_distance > 0 && detector_distance >= 0 && detector_width > 0,
                CTR_ERR_INVALID_ARGUMENT, "ctr_fdk_weight: bad geometry");
    long long total = (long long)nu * nv * n_views;
    if (total == 0) return CTR_OK;
    fdk_weight_kernel<<<grid1d(total, 256), 256, 0, as_stream(stream)>>>(
        proj, layout, nu, nv, n_views, source_distance + detector_distance,
        detector_width / nu);
    CTR_CHECK_LAUNCH();
    return CTR_OK;
}

int ctr_log_normalize(void *stream, float *data, size_t count, float base, float floor_value)
{
    clear_error();
    CTR_REQUIRE(data || count == 0, CTR_ERR_NULL_POINTER, "ctr_log_normalize: NULL buffer");
    CTR_REQUIRE(base > 0, CTR_ERR_INVALID_ARGUMENT, "ctr_log_normalize: base must be > 0");
    if (count == 0) return CTR_OK;
    log_normalize_kernel<<<grid1d((long long)count, 256), 256, 0, as_stream(stream)>>>(
        data, (long long)count, base, floor_value);
    CTR_CHECK_LAUNCH();
    return CTR_OK;
}

int ctr_clip_negative(void *stream, float *data, size_t count)
{
    clear_error();
    CTR_REQUIRE(data || count == 0, CTR_ERR_NULL_POINTER, "ctr_clip_negative: NULL buffer");
    if (count == 0) return CTR_OK;
    clip_kernel<<<grid1d((long long)count, 256), 256, 0, as_stream(stream)>>>(data,
                                                                             (long long)count);
    CTR_CHECK_LAUNCH();
    return CTR_OK;
}

int ctr_fill(void *stream, float *data, size_t count, float value)
{
    clear_error();
    CTR_REQUIRE(data || count == 0, CTR_ERR_NULL_POINTER, "ctr_fill: NULL buffer");
    if (count == 0) return CTR_OK;
    fill_kernel<<<grid1d((long long)count, 256), 256, 0, as_stream(stream)>>>(
        data, (long long)count, value);
    CTR_CHECK_LAUNCH();
    return CTR_OK;
}

int ctr_render_ellipsoids(void *stream, ctr_workspace *ws, const double *ellipsoids, int count,
                          float *volume, int X, int Y, int Z, double volume_width)
{
    clear_error();
    CTR_REQUIRE(volume && (ellipsoids || count == 0), CTR_ERR_NULL_POINTER,
                "[CTAnalyzer Error] ctr_render_ellipsoids: NULL buffer");
    CTR_REQUIRE(count >= 0 && count <= 4096, CTR_ERR_INVALID_ARGUMENT,
                "ctr_render_ellipsoids: count %d outside [0, 4096]", count);
    CTR_REQUIRE(X > 0 && Y > 0 && Z > 0 && volume_width > 0, CTR_ERR_INVALID_ARGUMENT,
                "ctr_render_ellipsoids: bad volume");
    if (count == 0) return CTR_OK;
    std::vector<EllipsoidRec> recs(count);
    for (int e = 0; e < count; e++) {
        int rc = prepare_ellipsoid(ellipsoids + 10 * (size_t)e, e, &recs[e]);
        if (rc) return rc;
    }
    ctr_workspace *w = nullptr;
    int rc = workspace_resolve(ws, &w);
    if (rc) return rc;
    void *p = nullptr;
    if ((rc = workspace_get(w, SLOT_ELLIPSOIDS, sizeof(EllipsoidRec) * count, &p))) return rc;
    hipStream_t s = as_stream(stream);
    CTR_TRY_HIP(hipMemcpyAsync(p, recs.data(), sizeof(EllipsoidRec) * count,
                               hipMemcpyHostToDevice, s));
    long long total = (long long)X * Y * Z;
    render_kernel<<<grid1d(total, 256), 256, 0, s>>>((const EllipsoidRec *)p, count, volume, X,
                                                     Y, Z, volume_width / X);
    CTR_CHECK_LAUNCH();
    // the host records must outlive the copy
    CTR_TRY_HIP(hipStreamSynchronize(s));
    return CTR_OK;
}

int ctr_project_ellipsoids(void *stream, int geometry, const double *ellipsoids, int count,
                           double source_distance, double detector_distance,
                           double detector_width, double rotation_count, double rotation_offset,
                           float *out, int layout, int nu, int nv, int n_views, int accumulate)
{
    clear_error();
    CTR_REQUIRE(out && (ellipsoids || count == 0), CTR_ERR_NULL_POINTER,
                "ctr_project_ellipsoids: NULL buffer");
    CTR_REQUIRE(count >= 0 && count <= 4096, CTR_ERR_INVALID_ARGUMENT,
                "ctr_project_ellipsoids: count %d outside [0, 4096]", count);
    CTR_REQUIRE(geometry == CTR_GEOMETRY_CONE || geometry == CTR_GEOMETRY_PARALLEL,
                CTR_ERR_INVALID_ARGUMENT, "unknown geometry %d", geometry);
    CTR_REQUIRE(layout == CTR_LAYOUT_PROJECTIONS || layout == CTR_LAYOUT_SINOGRAMS,
                CTR_ERR_INVALID_ARGUMENT, "unknown layout %d", layout);
    CTR_REQUIRE(nu > 0 && nv > 0 && n_views > 0 && detector_width > 0,
                CTR_ERR_INVALID_ARGUMENT, "ctr_project_ellipsoids: bad detector");
    CTR_REQUIRE(geometry == CTR_GEOMETRY_PARALLEL || (source_distance > 0 && detector_distance >= 0),
                CTR_ERR_INVALID_ARGUMENT, "ctr_project_ellipsoids: bad cone geometry");
    const long long blocks = (long long)((nu + kEllTile - 1) / kEllTile) *
                             ((nv + kEllTile - 1) / kEllTile) * n_views;
    CTR_REQUIRE(blocks < (1LL << 31), CTR_ERR_INVALID_ARGUMENT, "too many rays");
    RayGeometry G;
    G.geometry = geometry;
    G.layout = layout;
    G.nu = nu;
    G.nv = nv;
    G.n_views = n_views;
    G.sod = source_distance;
    G.odd = detector_distance;
    G.delta = detector_width / nu;
    G.rot = rotation_count;
    G.offset = rotation_offset;
    G.out = out;
    hipStream_t s = as_stream(stream);
    // groups of kEllGroup ellipsoids, each launch adding to the previous
    // ones; with none, one launch still clears `out` unless accumulating
    const int groups = count > 0 ? (count + kEllGroup - 1) / kEllGroup : (accumulate ? 0 : 1);
    for (int gi = 0; gi < groups; gi++) {
        const int e0 = gi * kEllGroup;
        EllipsoidGroup E;
        E.count = count - e0 < kEllGroup ? count - e0 : kEllGroup;
        for (int q = 0; q < E.count; q++) {
            const double *P = ellipsoids + 10 * (size_t)(e0 + q);
            EllipsoidRec R;
            int rc = prepare_ellipsoid(P, e0 + q, &R);
            if (rc) return rc;
            for (int a = 0; a < 3; a++) E.c[q][a] = R.c[a];
            for (int a = 0; a < 9; a++) E.m[q][a] = R.m[a];
            E.density[q] = R.density;
            E.radius[q] = fmax(P[3], fmax(P[4], P[5]));
        }
        G.accumulate = accumulate || gi > 0;
        project_ellipsoids_kernel<<<(unsigned)blocks, 256, 0, s>>>(E, G);
        CTR_CHECK_LAUNCH();
    }
    return CTR_OK;
}

int ctr_poisson_log(void *stream, float *data, size_t count, double i0, uint64_t seed)
{
    clear_error();
    CTR_REQUIRE(data || count == 0, CTR_ERR_NULL_POINTER, "ctr_poisson_log: NULL buffer");
    CTR_REQUIRE(i0 > 0, CTR_ERR_INVALID_ARGUMENT, "ctr_poisson_log: i0 must be > 0");
    if (count == 0) return CTR_OK;
    poisson_log_kernel<<<grid1d((long long)count, 256), 256, 0, as_stream(stream)>>>(
        data, (long long)count, i0, (unsigned long long)seed);
    CTR_CHECK_LAUNCH();
    return CTR_OK;
}

}  // extern "C"
