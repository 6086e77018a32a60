// ctr_internal.h -- shared internals of the gfx950 reconstruction library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "ctrecon.h"

namespace ctr {

// Records the calling thread's error text and returns `code`.
int fail(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
void clear_error();

inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

#define CTR_TRY_HIP(call)                                                        \
    do {                                                                         \
        hipError_t ctr_e_ = (call);                                              \
        if (ctr_e_ != hipSuccess)                                                \
            return ::ctr::fail(CTR_ERR_HIP, "%s failed: %s", #call,              \
                               hipGetErrorString(ctr_e_));                       \
    } while (0)

#define CTR_CHECK_LAUNCH() CTR_TRY_HIP(hipGetLastError())

#define CTR_REQUIRE(cond, code, ...)                                             \
    do {                                                                         \
        if (!(cond)) return ::ctr::fail((code), __VA_ARGS__);                    \
    } while (0)

// Workspace slots.  Each slot is one device buffer grown on demand.
enum Slot {
    SLOT_GEOMETRY = 0,   // per-view projective rows, 3 float4 per view
    SLOT_VIEW_MAX = 1,   // per-row max |q| of the filter output (16-bit staging)
    SLOT_STAGE_F32 = 2,  // fp32 staged projections before rounding to 16 bit
    SLOT_ELLIPSOIDS = 3, // ellipsoid records for the phantom renderer
    SLOT_FOOTPRINT = 4,  // per (tile column, view) detector footprint bounds
    SLOT_REDUCE = 5,     // min / max reduction keys
    SLOT_VOLUME_T = 6,   // forward projector: the volume with z fastest
    SLOT_COUNT = 7
};

// Resolves NULL to the per-device default workspace.
int workspace_resolve(ctr_workspace *ws, ctr_workspace **out);
// Returns a device buffer of at least `bytes` bytes for `slot`.
int workspace_get(ctr_workspace *ws, int slot, size_t bytes, void **ptr);
// Current value of a CTR_OPTION_* of a resolved workspace.
int workspace_option(const ctr_workspace *ws, int option);
// The load-time value of a CTR_OPTION_* (the default workspaces' value).
int default_option(int option);

inline bool is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }
inline int ceil_pow2(int n)
{
    int p = 1;
    while (p < n) p <<= 1;
    return p;
}
inline int log2_exact(int n)
{
    int l = 0;
    while ((1 << l) < n) l++;
    return l;
}

// Staged texel formats: CTR_F32, or the 16-bit value-and-step texels of
// CTR_F16 / CTR_BF16, which carry a per-view scale.
constexpr bool staged_dtype_valid(int d) { return d == CTR_F32 || d == CTR_F16 || d == CTR_BF16; }
constexpr bool staged_is_16bit(int d) { return d == CTR_F16 || d == CTR_BF16; }

// Layout of the staged (zero-bordered, view-major) projections.
struct StagedLayout {
    int pitch_u;   // texels per staged row
    int pitch_v;   // staged rows per view
    size_t view;   // texels per view
    size_t bytes;  // total bytes
};
StagedLayout staged_layout(int nu, int nv, int n_views, int dtype);

// Workspace bytes of per-view geometry: 3 float4 (FP32 rows for tile
// footprints) + 12 doubles (FP64 rows for voxel coordinates).
constexpr size_t kGeometryBytesPerView = 3 * 16 + 12 * 8;

// Per-view projective rows in voxel-index space (built in FP64 on device).
// rec[3n+0] = (U_i, U_k, U_0, V_j), rec[3n+1] = (V_i, V_k, V_0, 0),
// rec[3n+2] = (D_i, D_k, D_0, 0):
//   den = D.(i,k,1) = (SOD - depth)/SOD   (1 for parallel beams)
//   fu  = U.(i,k,1) / den,  fv = (V.(i,k,1) + V_j j) / den
int build_geometry(hipStream_t s, ctr_workspace *ws, int geometry, int n_views,
                   double sod, double odd, double detector_width,
                   double volume_width, double rotation_count,
                   double rotation_offset, int X, int Y, int Z, int nu, int nv,
                   const float4 **out);

int check_geometry(int geometry, double sod, double odd, double detector_width,
                   double volume_width, int X, int Y, int Z, int nu, int nv,
                   int n_views);

// Workgroups are dispatched to the 8 XCDs round-robin; this renumbering of a
// 1D grid hands each XCD one contiguous range of block indices, so blocks
// that read neighbouring data share that XCD's L2.
__device__ inline int xcd_tile_order(int b, int nblocks)
{
    const int xcd = b & 7, k = b >> 3;
    const int q = nblocks >> 3, r = nblocks & 7;
    return xcd * q + min(xcd, r) + k;
}

}  // namespace ctr
