"""ctrecon -- Python binding of the gfx950 CT reconstruction library.

The library itself is ``libctrecon.so`` (HIP kernels + the C ABI declared in
``include/ctrecon.h`` + the C++ drop-in classes FBP2D / FBP3D / FFFT).  This
module binds the C ABI with ctypes; torch is used only to own device memory
and to name the current HIP stream.  Every call goes to the HIP library: if
``libctrecon.so`` is missing, ``lib()`` raises instead of falling back.

The directory name is not a Python identifier, so callers load it with
``load()`` from ``__graft_entry__`` or by path (see tests/conftest.py).
"""
import ctypes
import math
import os
import subprocess
from dataclasses import dataclass, replace

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libctrecon.so")

CTR_OK = 0
F32, F16, BF16 = 0, 1, 2
OPTION_BACKPROJECT_KERNEL, OPTION_FORWARD_KERNEL, OPTION_FORWARD_ADDRESSING = 0, 1, 2
KERNEL_AUTO, KERNEL_LDS, KERNEL_GATHER, KERNEL_RAY, KERNEL_POINTER = 0, 1, 2, 3, 4
RAM_LAK, SHEPP_LOGAN, COSINE = 0, 1, 2
CONE, PARALLEL = 0, 1
PROJECTIONS, SINOGRAMS = 0, 1
BLIT_REAL_TO_COMPLEX, BLIT_COMPLEX_TO_COMPLEX = 0, 1
BLIT_COMPLEX_TO_REAL_R, BLIT_COMPLEX_TO_REAL_I = 2, 3

_LIB = None


class CtrError(RuntimeError):
    """A ctr_* call returned a negative status."""

    def __init__(self, code, message):
        super().__init__(f"ctrecon error {code}: {message}")
        self.code = code


def build(jobs=8):
    """Compile libctrecon.so for gfx950 (always recompiles)."""
    subprocess.run(["make", "-B", "-j", str(jobs), "-C", HERE], check=True)


_v, _i, _u, _d, _f = (ctypes.c_void_p, ctypes.c_int, ctypes.c_uint64,
                      ctypes.c_double, ctypes.c_float)
_sz = ctypes.c_size_t

# name -> argument types (all return int status unless listed in _RESTYPE)
SIGNATURES = {
    "ctr_last_error": [],
    "ctr_version": [],
    "ctr_synchronize": [],
    "ctr_workspace_create": [_v],
    "ctr_workspace_destroy": [_v],
    "ctr_workspace_reserve": [_v, _i, _i],
    "ctr_workspace_set_option": [_v, _i, _i],
    "ctr_workspace_get_option": [_v, _i],
    "ctr_fft_radix2": [_v, _v, _i, _i],
    "ctr_inverse_fft_radix2": [_v, _v, _i, _i],
    "ctr_fft_shift": [_v, _v, _v, _i, _i, _i],
    "ctr_blit": [_v, _i, _v, _i, _i, _v, _i, _i],
    "ctr_conjugate": [_v, _v, _sz],
    "ctr_scale_complex": [_v, _v, _sz, _f],
    "ctr_fft_bit_reverse": [_v, _v, _i, _i],
    "ctr_fft_single_stage": [_v, _v, _v, _i, _i, _i],
    "ctr_scale_real": [_v, _v, _sz, _f],
    "ctr_apply_filter": [_v, _v, _i, _i, _i],
    "ctr_filter_rows": [_v, _v, _v, _i, _i, _i, _i],
    "ctr_filter_projections": [_v, _v, _v, _i, _i, _i, _i, _i, _i, _i, _d, _d,
                               _d, _v, _i, _v],
    "ctr_staged_layout": [_i, _i, _i, _i, _v, _v, _v],
    "ctr_stage_projections": [_v, _v, _v, _i, _i, _i, _i, _v, _i, _v],
    "ctr_backproject": [_v, _v, _i, _v, _i, _v, _i, _i, _i, _d, _d, _d, _d, _d,
                        _d, _v, _i, _i, _i, _f, _i, _i],
    "ctr_backproject_noalloc": [_v, _i, _v, _i, _v, _i, _i, _i, _d, _d, _d, _d, _d, _d,
                                _v, _i, _i, _i, _f, _i, _i],
    "ctr_forward_project": [_v, _v, _i, _v, _i, _i, _i, _d, _d, _d, _d, _d, _d,
                            _v, _i, _i, _i, _i, _d],
    "ctr_fdk_weight": [_v, _v, _i, _i, _i, _i, _d, _d, _d],
    "ctr_log_normalize": [_v, _v, _sz, _f, _f],
    "ctr_clip_negative": [_v, _v, _sz],
    "ctr_fill": [_v, _v, _sz, _f],
    "ctr_rebin_fan_to_parallel": [_v, _v, _i, _i, _d, _d, _d, _d, _v, _i, _d],
    "ctr_render_ellipsoids": [_v, _v, _v, _i, _v, _i, _i, _i, _d],
    "ctr_poisson_log": [_v, _v, _sz, _d, _u],
    "ctr_project_ellipsoids": [_v, _i, _v, _i, _d, _d, _d, _d, _d, _v, _i, _i, _i, _i, _i],
    "ctr_decode_frames": [_v, _v, _i, _i, _i, _i, _i, _v, _i, _i, _i, _i, _i],
    "ctr_min_max": [_v, _v, _v, _sz, _v],
    "ctr_normalize_range": [_v, _v, _sz, _v],
    "ctr_slices_unorm16": [_v, _v, _i, _i, _i, _i, _i, _v],
}
_RESTYPE = {"ctr_last_error": ctypes.c_char_p}


class BackprojectJob(ctypes.Structure):
    """ctr_backproject_job of include/ctrecon.h."""
    _fields_ = [("staged", _v), ("view_scale", _v), ("volume", _v),
                ("geometry", _i), ("nu", _i), ("nv", _i), ("n_views", _i),
                ("X", _i), ("Y", _i), ("Z", _i),
                ("source_distance", _d), ("detector_distance", _d),
                ("detector_width", _d), ("volume_width", _d),
                ("rotation_count", _d), ("rotation_offset", _d),
                ("scale", _f), ("accumulate", _i), ("clip_negative", _i)]


SIGNATURES["ctr_backproject_packed"] = [_v, _v, _i, ctypes.POINTER(BackprojectJob), _i]
MAX_PACKED = 16


def lib():
    """The loaded libctrecon.so; raises if the HIP library is not built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: the HIP library is not built "
                "(run __graft_entry__.build()); there is no CPU fallback")
        L = ctypes.CDLL(LIB_PATH)
        for name, args in SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = _RESTYPE.get(name, ctypes.c_int)
        _LIB = L
    return _LIB


def last_error():
    return lib().ctr_last_error().decode()


def check(rc):
    if rc != CTR_OK:
        raise CtrError(rc, last_error())


# ---------------------------------------------------------------------------
# torch plumbing
# ---------------------------------------------------------------------------

def _torch():
    import torch
    return torch


def stream_handle(stream=None):
    """hipStream_t of `stream` (default: torch's current stream) as int."""
    torch = _torch()
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def ptr(t):
    if t is None:
        return None
    return ctypes.c_void_p(t.data_ptr())


def require_device():
    torch = _torch()
    if not torch.cuda.is_available():
        raise RuntimeError("ctrecon needs a HIP device (torch.cuda.is_available() is False)")
    lib()


class Workspace:
    """Owns a ctr_workspace (per-view geometry tables and staging scratch)."""

    def __init__(self):
        h = ctypes.c_void_p()
        check(lib().ctr_workspace_create(ctypes.byref(h)))
        self.handle = h

    def reserve(self, n_views, max_fft_length=0):
        check(lib().ctr_workspace_reserve(self.handle, n_views, max_fft_length))

    def set_option(self, option, value):
        """ctr_workspace_set_option; returns the previous value."""
        rc = lib().ctr_workspace_set_option(self.handle, int(option), int(value))
        check(rc if rc < 0 else CTR_OK)
        return rc

    def get_option(self, option):
        rc = lib().ctr_workspace_get_option(self.handle, int(option))
        check(rc if rc < 0 else CTR_OK)
        return rc

    def close(self):
        if self.handle:
            check(lib().ctr_workspace_destroy(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------------------
# Geometry
# ---------------------------------------------------------------------------

@dataclass(frozen=True)
class Geometry:
    """Scan + grid description with the reference's parameter meaning.

    source_distance = ``xray_source_distance`` (SOD), detector_distance =
    ``detector_panel_distance`` (centre -> detector), detector_width and
    volume_width in the same length unit (FBP3D.h:45,58)."""
    kind: int
    nu: int
    nv: int
    n_views: int
    X: int
    Y: int
    Z: int
    detector_width: float
    volume_width: float
    source_distance: float = 0.0
    detector_distance: float = 0.0
    rotation_count: float = 1.0
    rotation_offset: float = 0.0

    @property
    def sdd(self):
        return self.source_distance + self.detector_distance

    @property
    def delta(self):
        return self.detector_width / self.nu

    @property
    def voxel(self):
        return self.volume_width / self.X

    @property
    def updates(self):
        return self.X * self.Y * self.Z * self.n_views

    def fbp_scale(self, fft_length):
        """Backprojection scale turning reference-convention filtered
        projections (gain n*g) into attenuation per unit length:
        (pi/N) / (n * delta), times SDD/SOD for cone beams."""
        s = math.pi / self.n_views / (fft_length * self.delta)
        if self.kind == CONE:
            s *= self.sdd / self.source_distance
        return s


def default_fft_length(width):
    return 1 << max(0, int(width - 1).bit_length())


def staged_layout(nu, nv, n_views, dtype=F32):
    pu, pv, nb = ctypes.c_int(), ctypes.c_int(), ctypes.c_size_t()
    check(lib().ctr_staged_layout(nu, nv, n_views, dtype, ctypes.byref(pu),
                                  ctypes.byref(pv), ctypes.byref(nb)))
    return pu.value, pv.value, nb.value


# ---------------------------------------------------------------------------
# Thin wrappers (torch tensors in, torch tensors out)
# ---------------------------------------------------------------------------

def _f32c(t):
    torch = _torch()
    assert t.is_cuda and t.dtype == torch.float32 and t.is_contiguous(), \
        "expected a contiguous float32 device tensor"
    return t


def fft_radix2(z, inverse=False, stream=None):
    """In-place row FFT of a complex64 tensor (rows, n)."""
    torch = _torch()
    assert z.dtype == torch.complex64 and z.is_contiguous()
    n = z.shape[-1]
    rows = z.numel() // n
    fn = lib().ctr_inverse_fft_radix2 if inverse else lib().ctr_fft_radix2
    check(fn(stream_handle(stream), ptr(z), n, rows))
    return z


def filter_rows(x, filter_type=RAM_LAK, fft_length=0, out=None, stream=None):
    _f32c(x)
    out = x if out is None else _f32c(out)
    width = x.shape[-1]
    check(lib().ctr_filter_rows(stream_handle(stream), ptr(x), ptr(out), width,
                                x.numel() // width, fft_length, filter_type))
    return out


def fdk_weight(proj, geom, layout=PROJECTIONS, stream=None):
    _f32c(proj)
    check(lib().ctr_fdk_weight(stream_handle(stream), ptr(proj), layout, geom.nu,
                               geom.nv, geom.n_views, geom.source_distance,
                               geom.detector_distance, geom.detector_width))
    return proj


def stage_projections(proj, geom, dtype=F32, layout=PROJECTIONS, ws=None,
                      stream=None):
    """Plain projections -> (staged uint8 buffer, view_scale or None)."""
    torch = _torch()
    _f32c(proj)
    _, _, nbytes = staged_layout(geom.nu, geom.nv, geom.n_views, dtype)
    staged = torch.empty(nbytes, dtype=torch.uint8, device=proj.device)
    vs = None if dtype == F32 else torch.empty(geom.n_views, dtype=torch.float32,
                                               device=proj.device)
    check(lib().ctr_stage_projections(stream_handle(stream),
                                      ws.handle if ws else None, ptr(proj), layout,
                                      geom.nu, geom.nv, geom.n_views, ptr(staged),
                                      dtype, ptr(vs)))
    return staged, vs


def filter_projections(proj, geom, filter_type=RAM_LAK, fft_length=0, dtype=F32,
                       fdk=True, layout=PROJECTIONS, ws=None, staged=None,
                       view_scale=None, stream=None):
    """FDK weight + ramp filter straight into the staged layout."""
    torch = _torch()
    _f32c(proj)
    _, _, nbytes = staged_layout(geom.nu, geom.nv, geom.n_views, dtype)
    if staged is None:
        staged = torch.empty(nbytes, dtype=torch.uint8, device=proj.device)
    if view_scale is None and dtype != F32:
        view_scale = torch.empty(geom.n_views, dtype=torch.float32, device=proj.device)
    check(lib().ctr_filter_projections(
        stream_handle(stream), ws.handle if ws else None, ptr(proj), layout, geom.nu,
        geom.nv, geom.n_views, fft_length, filter_type, 1 if fdk else 0,
        geom.source_distance, geom.detector_distance, geom.detector_width,
        ptr(staged), dtype, ptr(view_scale)))
    return staged, view_scale


def backproject(staged, geom, dtype=F32, view_scale=None, scale=1.0, volume=None,
                accumulate=False, clip=False, ws=None, stream=None):
    torch = _torch()
    if volume is None:
        volume = torch.empty((geom.Z, geom.Y, geom.X), dtype=torch.float32,
                             device=staged.device)
    check(lib().ctr_backproject(
        stream_handle(stream), ws.handle if ws else None, geom.kind, ptr(staged),
        dtype, ptr(view_scale), geom.nu, geom.nv, geom.n_views,
        geom.source_distance, geom.detector_distance, geom.detector_width,
        geom.volume_width, geom.rotation_count, geom.rotation_offset, ptr(volume),
        geom.X, geom.Y, geom.Z, float(scale), 1 if accumulate else 0,
        1 if clip else 0))
    return volume


def default_option(option):
    """Option value of the current device's default workspace (fixed at
    library load from the environment)."""
    rc = lib().ctr_workspace_get_option(None, int(option))
    check(rc if rc < 0 else CTR_OK)
    return rc


def backproject_packed(items, dtype=F32, ws=None, stream=None):
    """One packed launch over up to 16 reconstructions.

    items: list of (staged, geom, view_scale, scale, volume, clip); returns
    the volumes."""
    if not 1 <= len(items) <= MAX_PACKED:
        raise ValueError(f"pack 1..{MAX_PACKED} jobs, got {len(items)}")
    jobs = (BackprojectJob * len(items))()
    for q, (staged, g, vs, scale, vol, clip) in enumerate(items):
        jobs[q] = BackprojectJob(
            staged.data_ptr(), vs.data_ptr() if vs is not None else None, vol.data_ptr(),
            g.kind, g.nu, g.nv, g.n_views, g.X, g.Y, g.Z, g.source_distance,
            g.detector_distance, g.detector_width, g.volume_width, g.rotation_count,
            g.rotation_offset, float(scale), 0, 1 if clip else 0)
    check(lib().ctr_backproject_packed(stream_handle(stream), ws.handle if ws else None,
                                       dtype, jobs, len(items)))
    return [it[4] for it in items]


def forward_project(volume, geom, layout=PROJECTIONS, step=1.0, out=None, ws=None,
                    stream=None):
    torch = _torch()
    _f32c(volume)
    if out is None:
        shape = ((geom.n_views, geom.nv, geom.nu) if layout == PROJECTIONS
                 else (geom.nv, geom.n_views, geom.nu))
        out = torch.empty(shape, dtype=torch.float32, device=volume.device)
    check(lib().ctr_forward_project(
        stream_handle(stream), ws.handle if ws else None, geom.kind, ptr(volume), geom.X,
        geom.Y, geom.Z,
        geom.source_distance, geom.detector_distance, geom.detector_width,
        geom.volume_width, geom.rotation_count, geom.rotation_offset, ptr(out),
        layout, geom.nu, geom.nv, geom.n_views, float(step)))
    return out


def render_ellipsoids(ellipsoids, volume, volume_width, ws=None, stream=None):
    e = np.ascontiguousarray(ellipsoids, dtype=np.float64).reshape(-1, 10)
    Z, Y, X = volume.shape
    check(lib().ctr_render_ellipsoids(
        stream_handle(stream), ws.handle if ws else None,
        e.ctypes.data_as(ctypes.c_void_p), e.shape[0], ptr(volume), X, Y, Z,
        float(volume_width)))
    return volume


def project_ellipsoids(ellipsoids, geom, layout=PROJECTIONS, out=None, accumulate=False,
                       device=None, stream=None):
    """Exact line integrals of ellipsoid records (ctr_project_ellipsoids)
    over the detector rays of `geom`; returns `out`, a float32 (N, nv, nu)
    or (nv, N, nu) device tensor."""
    torch = _torch()
    e = np.ascontiguousarray(ellipsoids, dtype=np.float64).reshape(-1, 10)
    if out is None:
        shape = ((geom.n_views, geom.nv, geom.nu) if layout == PROJECTIONS
                 else (geom.nv, geom.n_views, geom.nu))
        out = torch.empty(shape, dtype=torch.float32, device=device or "cuda")
    _f32c(out)
    check(lib().ctr_project_ellipsoids(
        stream_handle(stream), geom.kind, e.ctypes.data_as(ctypes.c_void_p), e.shape[0],
        geom.source_distance, geom.detector_distance, geom.detector_width,
        geom.rotation_count, geom.rotation_offset, ptr(out), layout, geom.nu, geom.nv,
        geom.n_views, 1 if accumulate else 0))
    return out


def poisson_log(proj, i0, seed, stream=None):
    _f32c(proj)
    check(lib().ctr_poisson_log(stream_handle(stream), ptr(proj), proj.numel(),
                                float(i0), int(seed) & (2**64 - 1)))
    return proj


def log_normalize(data, base, floor_value=1.1754944e-38, stream=None):
    _f32c(data)
    check(lib().ctr_log_normalize(stream_handle(stream), ptr(data), data.numel(),
                                  float(base), float(floor_value)))
    return data


def clip_negative(data, stream=None):
    _f32c(data)
    check(lib().ctr_clip_negative(stream_handle(stream), ptr(data), data.numel()))
    return data


def decode_frames(raw, file_w, file_h, channels, bytes_per_channel, out, target_w, target_h,
                  first_view=0, layout=PROJECTIONS, stream=None):
    """Raw frames (a contiguous uint8 device tensor of n_frames frames) ->
    views first_view.. of `out`, a float32 (N, target_h, target_w) projection
    stack or (target_h, N, target_w) sinogram stack (ctr_decode_frames)."""
    _f32c(out)
    frame_bytes = file_w * file_h * channels * bytes_per_channel
    assert raw.is_contiguous() and raw.numel() % frame_bytes == 0
    n_frames = raw.numel() // frame_bytes
    n_views = out.shape[0] if layout == PROJECTIONS else out.shape[1]
    check(lib().ctr_decode_frames(stream_handle(stream), ptr(raw), file_w, file_h, channels,
                                  bytes_per_channel, n_frames, ptr(out), layout, target_w,
                                  target_h, n_views, first_view))
    return out


def min_max(data, ws=None, stream=None):
    """(min, max) of a float32 device tensor as a 2-element device tensor."""
    torch = _torch()
    _f32c(data)
    out = torch.empty(2, dtype=torch.float32, device=data.device)
    check(lib().ctr_min_max(stream_handle(stream), ws.handle if ws else None, ptr(data),
                            data.numel(), ptr(out)))
    return out


def normalize_range(data, min_max_tensor, stream=None):
    _f32c(data)
    check(lib().ctr_normalize_range(stream_handle(stream), ptr(data), data.numel(),
                                    ptr(min_max_tensor)))
    return data


def slices_unorm16(volume, y0=0, count=None, stream=None):
    """y-slices of a (Z, Y, X) volume as unorm16 values, returned as an
    int16 device tensor (count, Z, X) holding the uint16 bit patterns."""
    torch = _torch()
    _f32c(volume)
    Z, Y, X = volume.shape
    count = Y - y0 if count is None else count
    out = torch.empty((count, Z, X), dtype=torch.int16, device=volume.device)
    check(lib().ctr_slices_unorm16(stream_handle(stream), ptr(volume), X, Y, Z, y0, count,
                                   ptr(out)))
    return out


def rebin_fan_to_parallel(fan, geom, nu_par, parallel_width, stream=None):
    """fan sinogram (N, nu_fan) -> parallel sinogram (N, nu_par)."""
    torch = _torch()
    _f32c(fan)
    n, nu_fan = fan.shape
    out = torch.empty((n, nu_par), dtype=torch.float32, device=fan.device)
    check(lib().ctr_rebin_fan_to_parallel(
        stream_handle(stream), ptr(fan), nu_fan, n, geom.source_distance,
        geom.detector_distance, geom.detector_width, geom.rotation_count,
        ptr(out), nu_par, float(parallel_width)))
    return out


# ---------------------------------------------------------------------------
# Seeded phantoms (benchmark inputs)
# ---------------------------------------------------------------------------

class SplitMix64:
    """Deterministic 64-bit generator for phantom parameters."""
    MASK = (1 << 64) - 1

    def __init__(self, seed):
        self.s = seed & self.MASK

    def next(self):
        self.s = (self.s + 0x9E3779B97F4A7C15) & self.MASK
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & self.MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & self.MASK
        return z ^ (z >> 31)

    def uniform(self, lo=0.0, hi=1.0):
        return lo + (hi - lo) * ((self.next() >> 11) * (1.0 / (1 << 53)))


def random_ellipsoids(seed, count, grid, pitch, semi_axes_vox, density=(0.1, 1.0),
                      planar=False, inner=0.8):
    """`count` ellipsoids (records of 10 doubles, world units) with centres
    uniform in the inner `inner` fraction of a grid of `grid` voxels of size
    `pitch`, semi-axes uniform in `semi_axes_vox` voxels, uniform random
    orientation and densities uniform in `density`.  planar=True makes 2D
    ellipses in the (x, z) plane (rotation about y only)."""
    g = SplitMix64(seed)
    X, Y, Z = grid
    out = np.zeros((count, 10))
    for e in range(count):
        cx = g.uniform(-0.5, 0.5) * inner * X * pitch
        cy = 0.0 if planar else g.uniform(-0.5, 0.5) * inner * Y * pitch
        cz = g.uniform(-0.5, 0.5) * inner * Z * pitch
        a = g.uniform(*semi_axes_vox) * pitch
        b = g.uniform(*semi_axes_vox) * pitch
        c = g.uniform(*semi_axes_vox) * pitch
        if planar:
            ang = (0.0, g.uniform(0.0, math.pi), 0.0)
            b = 1e6  # infinite extent along y: a 2D ellipse
        else:
            ang = (g.uniform(0.0, 2 * math.pi), g.uniform(0.0, 2 * math.pi),
                   g.uniform(0.0, 2 * math.pi))
        out[e] = (cx, cy, cz, a, b, c, *ang, g.uniform(*density))
    return out


# ---------------------------------------------------------------------------
# Benchmark / serving configurations
# ---------------------------------------------------------------------------

def config_geometry(name):
    """Geometry of a named workload.

    par2d_256   2D parallel, 256^2 image, 180 views over 180 deg, 367 bins
    fan2d_512   2D fan, 512^2 at 0.5 mm, 720 views / 360 deg, 768 bins x 1 mm,
                SOD 500, SDD 1000
    fdk_512     cone, 512^3 at 0.5 mm, 720 views / 360 deg, 768^2 x 0.8 mm,
                SOD 500, SDD 1000
    fdk_1440    cone, 1440 x 1470 x 1440 at 0.25 mm, 1440 views,
                1536^2 x 0.4 mm, SOD 500, SDD 1000
    fdk_256_s / fdk_512_s / fdk_1024_s  serving sizes
    """
    if name == "par2d_256":
        return Geometry(PARALLEL, 367, 1, 180, 256, 1, 256, 367.0, 256.0,
                        rotation_count=0.5)
    if name == "fan2d_512":
        return Geometry(CONE, 768, 1, 720, 512, 1, 512, 768.0, 256.0, 500.0, 500.0)
    if name == "fdk_512":
        return Geometry(CONE, 768, 768, 720, 512, 512, 512, 614.4, 256.0, 500.0, 500.0)
    if name == "fdk_1440":
        return Geometry(CONE, 1536, 1536, 1440, 1440, 1470, 1440, 614.4, 360.0,
                        500.0, 500.0)
    if name == "fdk_256_s":
        return Geometry(CONE, 384, 384, 360, 256, 256, 256, 614.4, 256.0, 500.0, 500.0)
    if name == "fdk_512_s":
        return Geometry(CONE, 768, 768, 720, 512, 512, 512, 614.4, 256.0, 500.0, 500.0)
    if name == "fdk_1024_s":
        return Geometry(CONE, 1024, 1024, 1200, 1024, 1024, 1024, 614.4, 256.0,
                        500.0, 500.0)
    raise KeyError(name)


def scaled_geometry(geom, factor, n_views=None):
    """Same scan shrunk by `factor` in every grid dimension (widths kept)."""
    return replace(geom, nu=max(1, geom.nu // factor),
                   nv=max(1, geom.nv // factor) if geom.nv > 1 else 1,
                   X=max(1, geom.X // factor),
                   Y=max(1, geom.Y // factor) if geom.Y > 1 else 1,
                   Z=max(1, geom.Z // factor),
                   n_views=n_views or max(1, geom.n_views // factor))
