"""CPU oracle for CTAnalyzer's reconstruction hot path -- TEST INFRASTRUCTURE.

This package is the checker.  Only tests/, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of bench.py import it; the product library never does.

It wraps two shared libraries built by ``oracle/Makefile``:

* ``liboracle.so`` -- the FP64 C restatement in ``ctr_oracle.c`` (each routine
  cites the reference file:line it follows);
* ``_ref/libreffft.so`` -- the reference's own CPU FFT
  (Source/FastFourierTransformCPU/FastFourierTransform.cpp) compiled from the
  reference checkout, used to pin the restatement.  It exists only where the
  reference checkout was present at build time.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None

RAM_LAK, SHEPP_LOGAN, COSINE = 0, 1, 2

_d = ctypes.c_double
_i = ctypes.c_int
_p = ctypes.c_void_p


def build():
    """Compile liboracle.so (and _ref/libreffft.so when the reference exists)."""
    import subprocess
    subprocess.run(["make", "-B", "-C", _HERE], check=True,
                   stdout=subprocess.DEVNULL)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        L.orc_reference_pi.restype = _d
        L.orc_fft_radix2.argtypes = [_p, _p, _i, _d]
        L.orc_inverse_fft_radix2.argtypes = [_p, _p, _i, _d]
        L.orc_fft_shift.argtypes = [_p, _p, _i, _i]
        L.orc_fft_bit_reverse.argtypes = [_p, _p, _i]
        L.orc_fft_single_stage.argtypes = [_p, _p, _p, _p, _i, _i, _d]
        L.orc_dft.argtypes = [_p, _p, _p, _p, _i, _i]
        L.orc_filter_gain.argtypes = [_i, _i, _i, _d]
        L.orc_filter_gain.restype = _d
        L.orc_filter_rows.argtypes = [_p, _i, _i, _i, _i, _d, _p, _i]
        L.orc_fdk_weight.argtypes = [_p, _i, _i, _i, _d, _d, _d]
        L.orc_log_normalize.argtypes = [_p, ctypes.c_longlong, _d, _d]
        L.orc_backproject_cone.argtypes = [_p, _i, _i, _i, _d, _d, _d, _d, _d,
                                           _d, _d, _i, _i, _i, _p, _i]
        L.orc_backproject_cone_region.argtypes = [_p, _i, _i, _i, _d, _d, _d,
                                                  _d, _d, _d, _d, _i, _i, _i,
                                                  _i, _i, _i, _i, _i, _i, _p,
                                                  _i]
        L.orc_backproject_parallel.argtypes = [_p, _i, _i, _i, _d, _d, _d, _d,
                                               _d, _i, _i, _i, _p, _i]
        L.orc_backproject_cone_cells.argtypes = [_p, _p, _i, _i, _i, _d, _d, _d,
                                                 _d, _d, _d, _d, _i, _i, _i, _p,
                                                 _i]
        L.orc_backproject_parallel_cells.argtypes = [_p, _p, _i, _i, _i, _d, _d,
                                                     _d, _d, _d, _i, _i, _i, _p,
                                                     _i]
        L.orc_backproject_cone_f32.argtypes = [_p, _i, _i, _i, _d, _d, _d, _d,
                                               _d, _d, _d, _i, _i, _i, _i, _i,
                                               _p, _i]
        L.orc_forward_cone.argtypes = [_p, _i, _i, _i, _d, _d, _d, _d, _i, _i,
                                       _i, _d, _d, _d, _p, _i]
        L.orc_forward_parallel.argtypes = [_p, _i, _i, _i, _d, _d, _i, _i, _i,
                                           _d, _d, _d, _p, _i]
        L.orc_rebin_fan_to_parallel.argtypes = [_p, _i, _i, _d, _d, _d, _d, _p,
                                                _i, _d]
        L.orc_ellipsoid_line_integral.argtypes = [_p, _i, _p, _p]
        L.orc_ellipsoid_line_integral.restype = _d
        L.orc_analytic_cone.argtypes = [_p, _i, _d, _d, _d, _i, _i, _i, _d, _d,
                                        _p, _i]
        L.orc_analytic_parallel.argtypes = [_p, _i, _d, _i, _i, _i, _d, _d, _p,
                                            _i]
        L.orc_render_ellipsoids.argtypes = [_p, _i, _i, _i, _i, _d, _p, _i]
        _LIB = L
    return _LIB


def ref_lib():
    """The reference's compiled CPU FFT, or None where it was not built."""
    global _REF
    if _REF is None:
        path = os.path.join(_HERE, "_ref", "libreffft.so")
        if not os.path.exists(path):
            return None
        L = ctypes.CDLL(path)
        L.ref_fft_radix2.argtypes = [_p, _p, _i]
        L.ref_inverse_fft_radix2.argtypes = [_p, _p, _i]
        L.ref_fft_shift.argtypes = [_p, _p, _i, _i]
        L.ref_pi.restype = _d
        _REF = L
    return _REF


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def threads_default():
    return os.cpu_count() or 1


REFERENCE_PI = None


def reference_pi():
    return lib().orc_reference_pi()


# --- FFT -------------------------------------------------------------------

def fft_radix2(z, pi=np.pi):
    z = np.asarray(z, dtype=np.complex128)
    re, im = _f64(z.real.copy()), _f64(z.imag.copy())
    if lib().orc_fft_radix2(_ptr(re), _ptr(im), re.size, pi) != 0:
        raise ValueError("length is not a power of 2")
    return re + 1j * im


def inverse_fft_radix2(z, pi=np.pi):
    z = np.asarray(z, dtype=np.complex128)
    re, im = _f64(z.real.copy()), _f64(z.imag.copy())
    if lib().orc_inverse_fft_radix2(_ptr(re), _ptr(im), re.size, pi) != 0:
        raise ValueError("length is not a power of 2")
    return re + 1j * im


def fft_shift(z, shift=None):
    z = np.asarray(z, dtype=np.complex128)
    n = z.size
    shift = n // 2 if shift is None else shift
    re, im = _f64(z.real.copy()), _f64(z.imag.copy())
    lib().orc_fft_shift(_ptr(re), _ptr(im), n, int(shift))
    return re + 1j * im


def fft_bit_reverse(z):
    """Bit-reversal permutation by swaps (the first pass of the FFT)."""
    z = np.asarray(z, dtype=np.complex128)
    re, im = _f64(z.real.copy()), _f64(z.imag.copy())
    if lib().orc_fft_bit_reverse(_ptr(re), _ptr(im), re.size) != 0:
        raise ValueError("length is not a power of 2")
    return re + 1j * im


def fft_single_stage(z, step_index, pi=np.pi):
    """One radix-2 stage of span 2 << step_index (read z, return the write)."""
    z = np.asarray(z, dtype=np.complex128)
    re, im = _f64(z.real.copy()), _f64(z.imag.copy())
    wre, wim = np.empty_like(re), np.empty_like(im)
    if lib().orc_fft_single_stage(_ptr(re), _ptr(im), _ptr(wre), _ptr(wim), re.size,
                                  int(step_index), pi) != 0:
        raise ValueError("bad length or stage")
    return wre + 1j * wim


def dft(z, sign=-1):
    z = np.asarray(z, dtype=np.complex128)
    re, im = _f64(z.real), _f64(z.imag)
    ore, oim = np.empty_like(re), np.empty_like(im)
    lib().orc_dft(_ptr(re), _ptr(im), _ptr(ore), _ptr(oim), re.size, sign)
    return ore + 1j * oim


def ref_fft_radix2(z):
    R = ref_lib()
    z = np.asarray(z, dtype=np.complex128)
    re, im = _f64(z.real.copy()), _f64(z.imag.copy())
    if R.ref_fft_radix2(_ptr(re), _ptr(im), re.size) != 0:
        raise ValueError("length is not a power of 2")
    return re + 1j * im


def ref_inverse_fft_radix2(z):
    R = ref_lib()
    z = np.asarray(z, dtype=np.complex128)
    re, im = _f64(z.real.copy()), _f64(z.imag.copy())
    if R.ref_inverse_fft_radix2(_ptr(re), _ptr(im), re.size) != 0:
        raise ValueError("length is not a power of 2")
    return re + 1j * im


def ref_fft_shift(z, shift):
    R = ref_lib()
    z = np.asarray(z, dtype=np.complex128)
    re, im = _f64(z.real.copy()), _f64(z.imag.copy())
    R.ref_fft_shift(_ptr(re), _ptr(im), re.size, int(shift))
    return re + 1j * im


# --- filter ----------------------------------------------------------------

def filter_gain(n, ftype, pi=np.pi):
    return np.array([lib().orc_filter_gain(i, n, ftype, pi) for i in range(n)])


def default_fft_length(width):
    """2**ceil(log2(width)) (FastFourierTransformGPU.cpp:258-280)."""
    return 1 << max(0, int(width - 1).bit_length())


def filter_rows(rows, ftype, npad=None, pi=np.pi, threads=None):
    rows = _f32(rows)
    width = rows.shape[-1]
    nrows = rows.size // width
    npad = default_fft_length(width) if npad is None else npad
    out = np.empty(rows.shape, dtype=np.float64)
    rc = lib().orc_filter_rows(_ptr(rows), width, nrows, npad, ftype, pi,
                               _ptr(out), threads or threads_default())
    if rc != 0:
        raise ValueError("bad filter arguments")
    return out


# --- weights ---------------------------------------------------------------

def fdk_weight(proj, sod, odd, detector_width):
    p = _f64(proj).copy()
    n, nv, nu = p.shape
    lib().orc_fdk_weight(_ptr(p), nu, nv, n, sod, odd, detector_width)
    return p


def log_normalize(proj, base, floor_value=0.0):
    p = _f64(proj).copy()
    lib().orc_log_normalize(_ptr(p), p.size, base, floor_value)
    return p


# --- backprojection --------------------------------------------------------

def backproject_cone(proj, sod, odd, detector_width, volume_width, rotation,
                     offset, scale, shape, threads=None):
    """proj (N, nv, nu) float32 -> volume (Z, Y, X) float64."""
    proj = _f32(proj)
    n, nv, nu = proj.shape
    X, Y, Z = shape
    vol = np.zeros((Z, Y, X), dtype=np.float64)
    lib().orc_backproject_cone(_ptr(proj), nu, nv, n, sod, odd, detector_width,
                               volume_width, rotation, offset, scale, X, Y, Z,
                               _ptr(vol), threads or threads_default())
    return vol


def backproject_cone_region(proj, sod, odd, detector_width, volume_width,
                            rotation, offset, scale, shape, box, threads=None):
    """FP64 cone backprojection of the voxel box ((i0,i1),(j0,j1),(k0,k1))
    of an (X, Y, Z) grid; returns (k1-k0, j1-j0, i1-i0)."""
    proj = _f32(proj)
    n, nv, nu = proj.shape
    X, Y, Z = shape
    (i0, i1), (j0, j1), (k0, k1) = box
    out = np.zeros((k1 - k0, j1 - j0, i1 - i0), dtype=np.float64)
    lib().orc_backproject_cone_region(_ptr(proj), nu, nv, n, sod, odd,
                                      detector_width, volume_width, rotation,
                                      offset, scale, X, Y, Z, i0, i1, j0, j1,
                                      k0, k1, _ptr(out),
                                      threads or threads_default())
    return out


def backproject_parallel(proj, detector_width, volume_width, rotation, offset,
                         scale, shape, threads=None):
    proj = _f32(proj)
    n, nv, nu = proj.shape
    X, Y, Z = shape
    vol = np.zeros((Z, Y, X), dtype=np.float64)
    lib().orc_backproject_parallel(_ptr(proj), nu, nv, n, detector_width,
                                   volume_width, rotation, offset, scale, X, Y,
                                   Z, _ptr(vol), threads or threads_default())
    return vol


def backproject_cone_cells(top, bot, sod, odd, detector_width, volume_width,
                           rotation, offset, scale, shape, threads=None):
    """Cone backprojection of views given per interpolation cell: top and
    bot (N, nv + 1, nu) hold, for cell c = iv + 1, detector row iv and row
    iv + 1 as that cell sees them (the dequantized 16-bit value-and-step
    texels).  Returns the (Z, Y, X) float64 volume."""
    top, bot = _f32(top), _f32(bot)
    assert top.shape == bot.shape
    n, cells, nu = top.shape
    X, Y, Z = shape
    vol = np.zeros((Z, Y, X), dtype=np.float64)
    lib().orc_backproject_cone_cells(_ptr(top), _ptr(bot), nu, cells - 1, n, sod,
                                     odd, detector_width, volume_width, rotation,
                                     offset, scale, X, Y, Z, _ptr(vol),
                                     threads or threads_default())
    return vol


def backproject_parallel_cells(top, bot, detector_width, volume_width,
                               rotation, offset, scale, shape, threads=None):
    """Parallel-beam counterpart of backproject_cone_cells."""
    top, bot = _f32(top), _f32(bot)
    assert top.shape == bot.shape
    n, cells, nu = top.shape
    X, Y, Z = shape
    vol = np.zeros((Z, Y, X), dtype=np.float64)
    lib().orc_backproject_parallel_cells(_ptr(top), _ptr(bot), nu, cells - 1, n,
                                         detector_width, volume_width, rotation,
                                         offset, scale, X, Y, Z, _ptr(vol),
                                         threads or threads_default())
    return vol


def backproject_cone_f32(proj, sod, odd, detector_width, volume_width,
                         rotation, offset, scale, shape, z_range=None,
                         threads=None):
    proj = _f32(proj)
    n, nv, nu = proj.shape
    X, Y, Z = shape
    z0, z1 = (0, Z) if z_range is None else z_range
    vol = np.zeros((z1 - z0, Y, X), dtype=np.float32)
    lib().orc_backproject_cone_f32(_ptr(proj), nu, nv, n, sod, odd,
                                   detector_width, volume_width, rotation,
                                   offset, scale, X, Y, Z, z0, z1, _ptr(vol),
                                   threads or threads_default())
    return vol


# --- forward projection ----------------------------------------------------

def forward_cone(vol, volume_width, sod, odd, detector_width, nu, nv, n_views,
                 rotation=1.0, offset=0.0, step=1.0, threads=None):
    vol = _f32(vol)
    Z, Y, X = vol.shape
    proj = np.zeros((n_views, nv, nu), dtype=np.float64)
    lib().orc_forward_cone(_ptr(vol), X, Y, Z, volume_width, sod, odd,
                           detector_width, nu, nv, n_views, rotation, offset,
                           step, _ptr(proj), threads or threads_default())
    return proj


def forward_parallel(vol, volume_width, detector_width, nu, nv, n_views,
                     rotation=0.5, offset=0.0, step=1.0, threads=None):
    vol = _f32(vol)
    Z, Y, X = vol.shape
    proj = np.zeros((n_views, nv, nu), dtype=np.float64)
    lib().orc_forward_parallel(_ptr(vol), X, Y, Z, volume_width,
                               detector_width, nu, nv, n_views, rotation,
                               offset, step, _ptr(proj),
                               threads or threads_default())
    return proj


def rebin_fan_to_parallel(fan, sod, odd, detector_width, rotation, nu_par,
                          parallel_width):
    """fan sinogram (N, nu_fan) -> parallel sinogram (N, nu_par), FP64."""
    fan = _f32(fan)
    n, nu_fan = fan.shape
    out = np.zeros((n, nu_par), dtype=np.float64)
    lib().orc_rebin_fan_to_parallel(_ptr(fan), nu_fan, n, sod, odd,
                                    detector_width, rotation, _ptr(out), nu_par,
                                    parallel_width)
    return out


# --- analytic phantoms -----------------------------------------------------

def ellipsoid_line_integral(ellipsoids, origin, direction):
    e = _f64(ellipsoids).reshape(-1, 10)
    o, d = _f64(origin), _f64(direction)
    return lib().orc_ellipsoid_line_integral(_ptr(e), e.shape[0], _ptr(o),
                                             _ptr(d))


def analytic_cone(ellipsoids, sod, odd, detector_width, nu, nv, n_views,
                  rotation=1.0, offset=0.0, threads=None):
    e = _f64(ellipsoids).reshape(-1, 10)
    proj = np.zeros((n_views, nv, nu), dtype=np.float64)
    lib().orc_analytic_cone(_ptr(e), e.shape[0], sod, odd, detector_width, nu,
                            nv, n_views, rotation, offset, _ptr(proj),
                            threads or threads_default())
    return proj


def analytic_parallel(ellipsoids, detector_width, nu, nv, n_views,
                      rotation=0.5, offset=0.0, threads=None):
    e = _f64(ellipsoids).reshape(-1, 10)
    proj = np.zeros((n_views, nv, nu), dtype=np.float64)
    lib().orc_analytic_parallel(_ptr(e), e.shape[0], detector_width, nu, nv,
                                n_views, rotation, offset, _ptr(proj),
                                threads or threads_default())
    return proj


def render_ellipsoids(ellipsoids, shape, volume_width, threads=None):
    e = _f64(ellipsoids).reshape(-1, 10)
    X, Y, Z = shape
    vol = np.zeros((Z, Y, X), dtype=np.float64)
    lib().orc_render_ellipsoids(_ptr(e), e.shape[0], X, Y, Z, volume_width,
                                _ptr(vol), threads or threads_default())
    return vol


# ---------------------------------------------------------------------------
# Formats either side of the path (projection ingest, slice output).  The
# reference does these with GraphicsCortex's Image class and GL conversions
# (FBP3D.cpp:74-195,780-826), which are absent: the restatement below follows
# the semantics the build documents (unorm conversion of channel 0, box
# filter when shrinking, linear between pixel centres when enlarging,
# unorm16 = round(clamp(v, 0, 1) * 65535)).  Parity with GraphicsCortex's
# resize and its Image flags is unpinned; the functions are pinned by closed
# forms in tests/test_oracle_formats.py.

def resample_matrix(n_dst, n_src):
    """(n_dst, n_src) weights of one axis: area coverage / scale when
    shrinking, linear between pixel centres (clamped) when enlarging."""
    m = np.zeros((n_dst, n_src), dtype=np.float64)
    if n_dst <= n_src:
        s = n_src / n_dst
        for d in range(n_dst):
            a, b = d * s, (d + 1) * s
            for i in range(int(np.floor(a)), min(int(np.ceil(b)), n_src)):
                m[d, i] = (min(b, i + 1) - max(a, i)) / s
    else:
        for d in range(n_dst):
            x = min(max((d + 0.5) * n_src / n_dst - 0.5, 0.0), n_src - 1)
            i0 = min(int(np.floor(x)), n_src - 1)
            f = x - i0
            if i0 + 1 < n_src and f > 0:
                m[d, i0] = 1.0 - f
                m[d, i0 + 1] = f
            else:
                m[d, i0] = 1.0
    return m


def decode_frame(raw, file_w, file_h, channels, bytes_per_channel, target_w, target_h):
    """One raw frame (bytes) -> float64 (target_h, target_w): channel 0 as
    unorm, resampled separably."""
    raw = np.frombuffer(bytes(raw), dtype=np.uint8)
    px = raw[:file_w * file_h * channels * bytes_per_channel].reshape(
        file_h, file_w, channels, bytes_per_channel)[:, :, 0, :]
    if bytes_per_channel == 1:
        img = px[..., 0].astype(np.float64) / 255.0
    else:
        img = (px[..., 0].astype(np.float64) + 256.0 * px[..., 1]) / 65535.0
    return resample_matrix(target_h, file_h) @ img @ resample_matrix(target_w, file_w).T


def unorm16(v):
    """round(clamp(v, 0, 1) * 65535) with the product in FP32 and
    round-half-even, NaN -> 0."""
    v = np.nan_to_num(np.asarray(v, dtype=np.float32), nan=0.0)
    c = np.clip(v, np.float32(0), np.float32(1))
    return np.rint(c * np.float32(65535)).astype(np.uint16)


def normalize_range(v):
    """(v - min) / (max - min), 0 when the range is empty (NaN ignored)."""
    v = np.asarray(v, dtype=np.float64)
    lo, hi = np.nanmin(v), np.nanmax(v)
    return np.zeros_like(v) if hi <= lo else (v - lo) / (hi - lo)
