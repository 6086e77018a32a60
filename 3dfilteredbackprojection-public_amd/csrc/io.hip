// io.hip -- the data formats on either side of the reconstruction path:
// raw scanner frames in (decode + resample, FBP3D::read_projections,
// FBP3D.cpp:74-119), 16-bit slices out (FBP3D::write_volume,
// FBP3D.cpp:178-195) and the volume min / max / range normalisation
// (FBP3D::compute_{min,max}_value_of_volume and normalize_histogram,
// FBP3D.cpp:780-826).
//
// The reference does this work with GraphicsCortex's Image class (file
// decode and resize on the host) and GL texture uploads / readbacks.  Here
// the host only moves bytes; decoding, resampling and the float <-> unorm16
// conversions run on the device:
//   * frames: channel 0 of each pixel, little-endian unsigned of 1 or 2
//     bytes, converted as an R8 / R16 texture upload converts them
//     (value / (2^(8 b) - 1));
//   * resampling, separable per axis: when shrinking, each target pixel is
//     the mean of the source interval it covers (box / area filter, partial
//     pixels weighted by coverage); when enlarging, linear interpolation
//     between source pixel centres, clamped at the edges;
//   * slices: unorm16 = round(clamp(v, 0, 1) * 65535), as a GL readback of
//     an R32F texture as UNSIGNED_SHORT.
#include "ctr_internal.h"

namespace ctr {

// Source taps of one target coordinate along an axis of `n_src` pixels
// resampled to `n_dst`: `count` taps from index `first`, weights by axis_w.
struct AxisTaps {
    int first, count;
    int area;      // 1: box filter over [a, b) in source pixels
    double a, b, s;
    float f;       // linear: weight of the second tap
};

__device__ inline AxisTaps axis_taps(int dst, int n_dst, int n_src)
{
    AxisTaps t;
    if (n_dst <= n_src) {
        t.area = 1;
        t.s = (double)n_src / (double)n_dst;
        t.a = dst * t.s;
        t.b = (dst + 1) * t.s;
        t.first = (int)floor(t.a);
        t.count = min((int)ceil(t.b), n_src) - t.first;
        t.f = 0.0f;
    } else {
        t.area = 0;
        t.a = t.b = t.s = 0.0;
        const double x = (dst + 0.5) * ((double)n_src / (double)n_dst) - 0.5;
        const double xc = fmin(fmax(x, 0.0), (double)(n_src - 1));
        t.first = min((int)floor(xc), n_src - 1);
        t.f = (float)(xc - t.first);
        t.count = t.first + 1 < n_src && t.f > 0.0f ? 2 : 1;
    }
    return t;
}

__device__ inline float axis_w(const AxisTaps &t, int q)
{
    if (t.area) {
        const double lo = fmax(t.a, (double)(t.first + q));
        const double hi = fmin(t.b, (double)(t.first + q + 1));
        return (float)((hi - lo) / t.s);
    }
    return t.count == 1 ? 1.0f : (q == 0 ? 1.0f - t.f : t.f);
}

__device__ inline float decode_px(const unsigned char *frame, size_t idx, int channels, int bpc)
{
    const unsigned char *p = frame + (size_t)idx * channels * bpc;
    if (bpc == 1) return (float)p[0] * (1.0f / 255.0f);
    const unsigned v = (unsigned)p[0] | ((unsigned)p[1] << 8);
    return (float)v * (1.0f / 65535.0f);
}

// One thread per target pixel of one frame (grid.y = frame).
__global__ void decode_frames_kernel(const unsigned char *raw, size_t frame_bytes, int file_w,
                                     int file_h, int channels, int bpc, float *out, int layout,
                                     int target_w, int target_h, int n_views, int first_view)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.z;
    const int f = blockIdx.y;
    if (x >= target_w) return;
    const unsigned char *frame = raw + (size_t)f * frame_bytes;
    const AxisTaps tx = axis_taps(x, target_w, file_w);
    const AxisTaps ty = axis_taps(y, target_h, file_h);
    float acc = 0.0f;
    for (int b = 0; b < ty.count; ++b) {
        const size_t row0 = (size_t)(ty.first + b) * file_w + tx.first;
        float row = 0.0f;
        for (int a = 0; a < tx.count; ++a)
            row = fmaf(axis_w(tx, a), decode_px(frame, row0 + a, channels, bpc), row);
        acc = fmaf(axis_w(ty, b), row, acc);
    }
    const int view = first_view + f;
    const size_t o = layout == CTR_LAYOUT_PROJECTIONS
                         ? ((size_t)view * target_h + y) * target_w + x
                         : ((size_t)y * n_views + view) * target_w + x;
    out[o] = acc;
}

// Order-preserving float <-> uint keys for atomic min / max of any sign.
__device__ inline unsigned float_key(float f)
{
    const unsigned u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

__device__ inline float key_float(unsigned k)
{
    return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
}

__global__ void minmax_init_kernel(unsigned *keys)
{
    keys[0] = 0xffffffffu;  // min
    keys[1] = 0u;           // max
}

__global__ __launch_bounds__(256) void minmax_kernel(const float *v, size_t count, unsigned *keys)
{
    unsigned lo = 0xffffffffu, hi = 0u;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
         i += (size_t)gridDim.x * blockDim.x) {
        const float x = v[i];
        if (x != x) continue;  // NaN
        const unsigned k = float_key(x);
        lo = min(lo, k);
        hi = max(hi, k);
    }
    for (int off = 32; off > 0; off >>= 1) {
        lo = min(lo, (unsigned)__shfl_xor((int)lo, off));
        hi = max(hi, (unsigned)__shfl_xor((int)hi, off));
    }
    if ((threadIdx.x & 63) == 0) {
        atomicMin(keys, lo);
        atomicMax(keys + 1, hi);
    }
}

__global__ void minmax_finish_kernel(const unsigned *keys, float *min_max)
{
    // no finite value: (0, 0)
    const bool empty = keys[0] == 0xffffffffu && keys[1] == 0u;
    min_max[0] = empty ? 0.0f : key_float(keys[0]);
    min_max[1] = empty ? 0.0f : key_float(keys[1]);
}

__global__ void normalize_range_kernel(float *v, size_t count, const float *min_max)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const float lo = min_max[0], range = min_max[1] - min_max[0];
    v[i] = range > 0.0f ? (v[i] - lo) / range : 0.0f;
}

__device__ inline unsigned short unorm16(float v)
{
    const float c = fminf(fmaxf(v, 0.0f), 1.0f);  // NaN -> 0
    return (unsigned short)__float2uint_rn(c * 65535.0f);
}

// y-slices y0 .. y0 + ny - 1 of a (Z, Y, X) volume as unorm16, each slice
// (Z rows of X), x fastest.
__global__ void slices_unorm16_kernel(const float *vol, int X, int Y, int Z, int y0, int ny,
                                      unsigned short *dst)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int z = blockIdx.y;
    const int s = blockIdx.z;
    if (x >= X) return;
    const float v = vol[((size_t)z * Y + (y0 + s)) * X + x];
    dst[((size_t)s * Z + z) * X + x] = unorm16(v);
}

}  // namespace ctr

using namespace ctr;

extern "C" {

int ctr_decode_frames(void *stream, const void *raw, int file_w, int file_h, int channels,
                      int bytes_per_channel, int n_frames, float *out, int layout, int target_w,
                      int target_h, int n_views, int first_view)
{
    clear_error();
    CTR_REQUIRE(raw && out, CTR_ERR_NULL_POINTER, "ctr_decode_frames: NULL buffer");
    CTR_REQUIRE(file_w > 0 && file_h > 0 && channels > 0 && target_w > 0 && target_h > 0 &&
                    n_frames >= 0 && n_views > 0,
                CTR_ERR_INVALID_ARGUMENT, "ctr_decode_frames: bad sizes");
    CTR_REQUIRE(bytes_per_channel == 1 || bytes_per_channel == 2, CTR_ERR_INVALID_ARGUMENT,
                "ctr_decode_frames: %d bytes per channel (1 or 2 supported)", bytes_per_channel);
    CTR_REQUIRE(layout == CTR_LAYOUT_PROJECTIONS || layout == CTR_LAYOUT_SINOGRAMS,
                CTR_ERR_INVALID_ARGUMENT, "unknown layout %d", layout);
    CTR_REQUIRE(first_view >= 0 && first_view + n_frames <= n_views, CTR_ERR_INVALID_ARGUMENT,
                "ctr_decode_frames: frames %d..%d outside %d views", first_view,
                first_view + n_frames - 1, n_views);
    CTR_REQUIRE(target_h <= 65535 && n_frames <= 65535, CTR_ERR_INVALID_ARGUMENT,
                "ctr_decode_frames: at most 65535 rows and frames per call");
    if (n_frames == 0) return CTR_OK;
    const size_t frame_bytes = (size_t)file_w * file_h * channels * bytes_per_channel;
    dim3 grid((target_w + 127) / 128, n_frames, target_h);
    decode_frames_kernel<<<grid, 128, 0, as_stream(stream)>>>(
        (const unsigned char *)raw, frame_bytes, file_w, file_h, channels, bytes_per_channel, out,
        layout, target_w, target_h, n_views, first_view);
    CTR_CHECK_LAUNCH();
    return CTR_OK;
}

int ctr_min_max(void *stream, ctr_workspace *ws, const float *data, size_t count,
                float *min_max)
{
    clear_error();
    CTR_REQUIRE(data && min_max, CTR_ERR_NULL_POINTER, "ctr_min_max: NULL buffer");
    ctr_workspace *w = nullptr;
    int rc = workspace_resolve(ws, &w);
    if (rc) return rc;
    void *keys = nullptr;
    if ((rc = workspace_get(w, SLOT_REDUCE, 2 * sizeof(unsigned), &keys))) return rc;
    hipStream_t s = as_stream(stream);
    minmax_init_kernel<<<1, 1, 0, s>>>((unsigned *)keys);
    CTR_CHECK_LAUNCH();
    const size_t blocks = count == 0 ? 1 : (count + 255) / 256 < 4096 ? (count + 255) / 256 : 4096;
    minmax_kernel<<<(unsigned)blocks, 256, 0, s>>>(data, count, (unsigned *)keys);
    CTR_CHECK_LAUNCH();
    minmax_finish_kernel<<<1, 1, 0, s>>>((const unsigned *)keys, min_max);
    CTR_CHECK_LAUNCH();
    return CTR_OK;
}

int ctr_normalize_range(void *stream, float *data, size_t count, const float *min_max)
{
    clear_error();
    CTR_REQUIRE(data && min_max, CTR_ERR_NULL_POINTER, "ctr_normalize_range: NULL buffer");
    if (count == 0) return CTR_OK;
    normalize_range_kernel<<<(unsigned)((count + 255) / 256), 256, 0, as_stream(stream)>>>(
        data, count, min_max);
    CTR_CHECK_LAUNCH();
    return CTR_OK;
}

int ctr_slices_unorm16(void *stream, const float *volume, int X, int Y, int Z, int y0,
                       int count, unsigned short *out)
{
    clear_error();
    CTR_REQUIRE(volume && out, CTR_ERR_NULL_POINTER, "ctr_slices_unorm16: NULL buffer");
    CTR_REQUIRE(X > 0 && Y > 0 && Z > 0 && Z <= 65535, CTR_ERR_INVALID_ARGUMENT,
                "ctr_slices_unorm16: bad volume (%d, %d, %d)", X, Y, Z);
    CTR_REQUIRE(y0 >= 0 && count >= 0 && y0 + count <= Y && count <= 65535,
                CTR_ERR_INVALID_ARGUMENT, "ctr_slices_unorm16: slices %d..%d outside [0, %d)",
                y0, y0 + count - 1, Y);
    if (count == 0) return CTR_OK;
    dim3 grid((X + 255) / 256, Z, count);
    slices_unorm16_kernel<<<grid, 256, 0, as_stream(stream)>>>(volume, X, Y, Z, y0, count, out);
    CTR_CHECK_LAUNCH();
    return CTR_OK;
}

}  // extern "C"
