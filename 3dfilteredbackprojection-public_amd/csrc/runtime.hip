// runtime.hip -- errors, workspaces and staged-layout bookkeeping.
#include <atomic>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>

#include "ctr_internal.h"

struct ctr_workspace {
    int device;
    void *buf[ctr::SLOT_COUNT];
    size_t cap[ctr::SLOT_COUNT];
    std::atomic<int> option[CTR_OPTION_COUNT];
    std::mutex lock;
};

namespace ctr {

// Initial option values of every workspace, from the environment as it is
// when the library is loaded (read once, here; never again per call).
struct OptionDefaults {
    int value[CTR_OPTION_COUNT];
    OptionDefaults()
    {
        value[CTR_OPTION_BACKPROJECT_KERNEL] =
            env_is("CTR_BACKPROJECT_KERNEL", "lds")      ? CTR_KERNEL_LDS
            : env_is("CTR_BACKPROJECT_KERNEL", "gather") ? CTR_KERNEL_GATHER
                                                         : CTR_KERNEL_AUTO;
        value[CTR_OPTION_FORWARD_KERNEL] =
            env_is("CTR_FORWARD_KERNEL", "ray") ? CTR_KERNEL_RAY : CTR_KERNEL_AUTO;
        value[CTR_OPTION_FORWARD_ADDRESSING] =
            env_is("CTR_FORWARD_ADDRESSING", "pointer") ? CTR_KERNEL_POINTER : CTR_KERNEL_AUTO;
    }
    static bool env_is(const char *name, const char *want)
    {
        const char *e = getenv(name);
        return e && !strcmp(e, want);
    }
};
static const OptionDefaults g_option_defaults;

static bool option_value_ok(int option, int value)
{
    switch (option) {
    case CTR_OPTION_BACKPROJECT_KERNEL:
        return value == CTR_KERNEL_AUTO || value == CTR_KERNEL_LDS || value == CTR_KERNEL_GATHER;
    case CTR_OPTION_FORWARD_KERNEL: return value == CTR_KERNEL_AUTO || value == CTR_KERNEL_RAY;
    case CTR_OPTION_FORWARD_ADDRESSING:
        return value == CTR_KERNEL_AUTO || value == CTR_KERNEL_POINTER;
    default: return false;
    }
}

int workspace_option(const ctr_workspace *ws, int option)
{
    return ws->option[option].load(std::memory_order_relaxed);
}

int default_option(int option) { return g_option_defaults.value[option]; }

static thread_local char g_error[512] = {0};

int fail(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_error, sizeof(g_error), fmt, ap);
    va_end(ap);
    return code;
}

void clear_error() { g_error[0] = 0; }

static std::mutex g_default_lock;
static ctr_workspace *g_default[64] = {nullptr};

int workspace_resolve(ctr_workspace *ws, ctr_workspace **out)
{
    if (ws) {
        *out = ws;
        return CTR_OK;
    }
    int dev = 0;
    CTR_TRY_HIP(hipGetDevice(&dev));
    CTR_REQUIRE(dev >= 0 && dev < 64, CTR_ERR_INVALID_ARGUMENT,
                "device index %d out of range", dev);
    std::lock_guard<std::mutex> g(g_default_lock);
    if (!g_default[dev]) {
        int rc = ctr_workspace_create(&g_default[dev]);
        if (rc != CTR_OK) return rc;
    }
    *out = g_default[dev];
    return CTR_OK;
}

int workspace_get(ctr_workspace *ws, int slot, size_t bytes, void **ptr)
{
    CTR_REQUIRE(slot >= 0 && slot < SLOT_COUNT, CTR_ERR_INVALID_ARGUMENT,
                "bad workspace slot %d", slot);
    std::lock_guard<std::mutex> g(ws->lock);
    if (ws->cap[slot] < bytes) {
        if (ws->buf[slot]) {
            CTR_TRY_HIP(hipDeviceSynchronize());
            CTR_TRY_HIP(hipFree(ws->buf[slot]));
            ws->buf[slot] = nullptr;
            ws->cap[slot] = 0;
        }
        size_t want = (bytes + 4095) & ~size_t(4095);
        hipError_t e = hipMalloc(&ws->buf[slot], want);
        if (e != hipSuccess) {
            ws->buf[slot] = nullptr;
            return fail(CTR_ERR_OUT_OF_MEMORY,
                        "workspace slot %d: hipMalloc(%zu) failed: %s", slot,
                        want, hipGetErrorString(e));
        }
        ws->cap[slot] = want;
    }
    *ptr = ws->buf[slot];
    return CTR_OK;
}

StagedLayout staged_layout(int nu, int nv, int n_views, int dtype)
{
    StagedLayout L;
    // one zero column before, at least two after; a multiple of 32 texels
    // (128 bytes) so that rows start on cache-line boundaries
    L.pitch_u = (nu + 3 + 31) & ~31;
    L.pitch_v = nv + 3;             // one zero row before, two after
    L.view = (size_t)L.pitch_u * (size_t)L.pitch_v;
    size_t texel = 4;               // fp32 value, or a 16-bit value and its step
    (void)dtype;
    L.bytes = L.view * (size_t)n_views * texel;
    return L;
}

}  // namespace ctr

extern "C" {

const char *ctr_last_error(void) { return ctr::g_error; }

int ctr_version(void) { return 10000; }

int ctr_synchronize(void)
{
    CTR_TRY_HIP(hipDeviceSynchronize());
    return CTR_OK;
}

int ctr_workspace_create(ctr_workspace **out)
{
    CTR_REQUIRE(out, CTR_ERR_NULL_POINTER, "ctr_workspace_create: out is NULL");
    int device = 0;
    CTR_TRY_HIP(hipGetDevice(&device));
    ctr_workspace *ws = new ctr_workspace();
    ws->device = device;
    for (int i = 0; i < ctr::SLOT_COUNT; i++) {
        ws->buf[i] = nullptr;
        ws->cap[i] = 0;
    }
    for (int o = 0; o < CTR_OPTION_COUNT; o++)
        ws->option[o].store(ctr::g_option_defaults.value[o]);
    *out = ws;
    return CTR_OK;
}

int ctr_workspace_set_option(ctr_workspace *ws, int option, int value)
{
    ctr::clear_error();
    CTR_REQUIRE(ws, CTR_ERR_INVALID_ARGUMENT,
                "ctr_workspace_set_option: the default workspace (NULL) is shared by every "
                "caller on the device; its options are fixed at library load");
    CTR_REQUIRE(option >= 0 && option < CTR_OPTION_COUNT, CTR_ERR_INVALID_ARGUMENT,
                "ctr_workspace_set_option: unknown option %d", option);
    CTR_REQUIRE(ctr::option_value_ok(option, value), CTR_ERR_INVALID_ARGUMENT,
                "ctr_workspace_set_option: value %d is not valid for option %d", value, option);
    return ws->option[option].exchange(value);
}

int ctr_workspace_get_option(ctr_workspace *ws, int option)
{
    ctr::clear_error();
    CTR_REQUIRE(option >= 0 && option < CTR_OPTION_COUNT, CTR_ERR_INVALID_ARGUMENT,
                "ctr_workspace_get_option: unknown option %d", option);
    // the default workspaces keep the load-time values: no device needed
    if (!ws) return ctr::default_option(option);
    return ctr::workspace_option(ws, option);
}

int ctr_workspace_destroy(ctr_workspace *ws)
{
    if (!ws) return CTR_OK;
    int cur = 0;
    CTR_TRY_HIP(hipGetDevice(&cur));
    CTR_TRY_HIP(hipSetDevice(ws->device));
    CTR_TRY_HIP(hipDeviceSynchronize());
    for (int i = 0; i < ctr::SLOT_COUNT; i++)
        if (ws->buf[i]) CTR_TRY_HIP(hipFree(ws->buf[i]));
    CTR_TRY_HIP(hipSetDevice(cur));
    delete ws;
    return CTR_OK;
}

int ctr_workspace_reserve(ctr_workspace *ws, int n_views, int max_fft_length)
{
    (void)max_fft_length;
    CTR_REQUIRE(n_views >= 0, CTR_ERR_INVALID_ARGUMENT, "n_views < 0");
    ctr_workspace *w = nullptr;
    int rc = ctr::workspace_resolve(ws, &w);
    if (rc) return rc;
    void *p = nullptr;
    if ((rc = ctr::workspace_get(w, ctr::SLOT_GEOMETRY,
                                 ctr::kGeometryBytesPerView * (size_t)(n_views + 1), &p)))
        return rc;
    return CTR_OK;
}

int ctr_staged_layout(int nu, int nv, int n_views, int dtype, int *pitch_u,
                      int *pitch_v, size_t *bytes)
{
    CTR_REQUIRE(nu > 0 && nv > 0 && n_views > 0, CTR_ERR_INVALID_ARGUMENT,
                "ctr_staged_layout: sizes must be positive (%d, %d, %d)", nu,
                nv, n_views);
    CTR_REQUIRE(ctr::staged_dtype_valid(dtype), CTR_ERR_INVALID_ARGUMENT,
                "unknown staged dtype %d", dtype);
    ctr::StagedLayout L = ctr::staged_layout(nu, nv, n_views, dtype);
    if (pitch_u) *pitch_u = L.pitch_u;
    if (pitch_v) *pitch_v = L.pitch_v;
    if (bytes) *bytes = L.bytes;
    return CTR_OK;
}

}  // extern "C"
