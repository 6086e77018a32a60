"""Every FFFT / FBP2D / FBP3D primitive of the C ABI against the oracle, on
the GPU, called through the C ABI itself (ctypes on libctrecon.so).

The FFFT primitives (FastFourierTransformGPU.cpp:93-256) move or scale data:
shifts, blits, conjugation, scaling and the bit-reversal permutation are held
bit-exact against the oracle / numpy.  The single radix-2 stage and the
Fourier-space filter do FP32 arithmetic and are held to relative errors of a
few FP32 ulps against the FP64 oracle.  The elementwise detector and volume
passes (FDK weight, log-normalisation, clipping, fill) and the ellipsoid
voxeliser are compared with the oracle on seeded inputs.
"""
import ctypes

import numpy as np
import pytest

from conftest import rel_max

pytestmark = pytest.mark.gpu


def _dev(dev, a, dtype=np.float32):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, dtype=dtype)).to(dev)


def _complex(rng, rows, n):
    return (rng.standard_normal((rows, n)) + 1j * rng.standard_normal((rows, n))).astype(
        np.complex64)


def _call(ctr, name, *args):
    ctr.check(getattr(ctr.lib(), name)(ctr.stream_handle(), *args))


def _host(t):
    import torch
    torch.cuda.synchronize()
    return t.cpu().numpy()


# --- FFFT::fft_shift / inverse_fft_shift (FastFourierTransformGPU.cpp:93-142)

@pytest.mark.parametrize("n", [1, 8, 12, 100, 1024, 2048, 8192, 16384])
def test_fft_shift_vs_oracle(ctr, orc, dev, n):
    """Out of place and in place (the in-place form rotates rows through LDS
    up to 8192 values and by three in-place reversals beyond), for +-n/2,
    odd shifts, a shift past n and no shift."""
    rng = np.random.default_rng(n)
    rows = 3
    z = _complex(rng, rows, n)
    for shift in sorted({n // 2, -(n // 2), 3, -5, n + 1, 0, 7 - n}):
        ref = np.stack([orc.fft_shift(r.astype(np.complex128), shift) for r in z])
        # the oracle's modulo is the true one; np.roll agrees for every n
        assert np.array_equal(ref, np.roll(z, shift, axis=1).astype(np.complex128))
        src = _dev(dev, z, np.complex64)
        dst = _dev(dev, np.zeros_like(z), np.complex64)
        _call(ctr, "ctr_fft_shift", ctr.ptr(src), ctr.ptr(dst), n, rows, shift)
        assert np.array_equal(_host(dst), ref.astype(np.complex64)), (n, shift)
        _call(ctr, "ctr_fft_shift", ctr.ptr(src), ctr.ptr(src), n, rows, shift)
        assert np.array_equal(_host(src), ref.astype(np.complex64)), ("in place", n, shift)


def test_fft_shift_matches_reference_golden(ctr, dev):
    """The reference's own compiled fft_shift output (golden vectors)."""
    import os
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "reference_fft.npz"),
                allow_pickle=False)
    for n in (8, 64, 512, 2048):
        z = g[f"fft_in_{n}"].astype(np.complex64)
        for key, shift in ((f"shift_pos_{n}", n // 2), (f"shift_neg_{n}", -(n // 2))):
            t = _dev(dev, z[None, :], np.complex64)
            _call(ctr, "ctr_fft_shift", ctr.ptr(t), ctr.ptr(t), n, 1, shift)
            assert np.array_equal(_host(t)[0], g[key].astype(np.complex64)), key


# --- FFFT::blit_texture_* (FastFourierTransformGPU.cpp:144-210) --------------

@pytest.mark.parametrize("mode", [0, 1, 2, 3])
@pytest.mark.parametrize("src_wh,dst_wh", [((5, 4), (8, 6)), ((9, 7), (4, 3)),
                                           ((6, 3), (4, 5)), ((16, 2), (16, 2))])
def test_blit_modes_pad_and_crop(ctr, dev, mode, src_wh, dst_wh):
    """The min(src, dst) extent is copied (pad when the target is larger,
    crop when it is smaller); target texels outside it keep their values."""
    rng = np.random.default_rng(mode * 100 + src_wh[0])
    (sw, sh), (dw, dh) = src_wh, dst_wh
    src_complex = mode in (1, 2, 3)
    dst_complex = mode in (0, 1)
    src = (_complex(rng, sh, sw) if src_complex
           else rng.standard_normal((sh, sw)).astype(np.float32))
    sentinel = (np.full((dh, dw), 7.0 - 3.0j, dtype=np.complex64) if dst_complex
                else np.full((dh, dw), 7.0, dtype=np.float32))
    w, h = min(sw, dw), min(sh, dh)
    ref = sentinel.copy()
    block = src[:h, :w]
    ref[:h, :w] = {0: block.astype(np.complex64), 1: block, 2: np.real(block),
                   3: np.imag(block)}[mode]
    ts = _dev(dev, src, np.complex64 if src_complex else np.float32)
    td = _dev(dev, sentinel, sentinel.dtype)
    _call(ctr, "ctr_blit", mode, ctr.ptr(ts), sw, sh, ctr.ptr(td), dw, dh)
    assert np.array_equal(_host(td), ref)


# --- conjugate / multiply / divide (FastFourierTransformGPU.cpp:212-236),
#     FBP2D::_int_texture_div (FBP2D.cpp:189-195) ----------------------------

def test_conjugate_and_scales_bit_exact(ctr, dev):
    rng = np.random.default_rng(9)
    z = _complex(rng, 7, 333)
    t = _dev(dev, z, np.complex64)
    _call(ctr, "ctr_conjugate", ctr.ptr(t), ctypes.c_size_t(z.size))
    assert np.array_equal(_host(t), np.conj(z))
    _call(ctr, "ctr_scale_complex", ctr.ptr(t), ctypes.c_size_t(z.size), ctypes.c_float(0.37))
    ref = np.conj(z)
    ref = (ref.real * np.float32(0.37)) + 1j * (ref.imag * np.float32(0.37))
    assert np.array_equal(_host(t), ref.astype(np.complex64))
    x = rng.standard_normal(1001).astype(np.float32)
    tx = _dev(dev, x)
    _call(ctr, "ctr_scale_real", ctr.ptr(tx), ctypes.c_size_t(x.size), ctypes.c_float(1 / 3))
    assert np.array_equal(_host(tx), x * np.float32(1 / 3))
    # count 0 touches nothing
    _call(ctr, "ctr_conjugate", ctr.ptr(t), ctypes.c_size_t(0))


# --- FFFT::reverse_bit_swap / fft_single_step (FastFourierTransformGPU.cpp:238-256)

@pytest.mark.parametrize("n", [1, 2, 16, 512, 2048, 8192])
def test_bit_reverse_vs_oracle(ctr, orc, dev, n):
    rng = np.random.default_rng(n + 1)
    z = _complex(rng, 4, n)
    t = _dev(dev, z, np.complex64)
    _call(ctr, "ctr_fft_bit_reverse", ctr.ptr(t), n, 4)
    ref = np.stack([orc.fft_bit_reverse(r) for r in z]).astype(np.complex64)
    assert np.array_equal(_host(t), ref)


@pytest.mark.parametrize("n", [2, 8, 256, 2048, 8192])
def test_single_stages_vs_oracle(ctr, orc, dev, n):
    """Every stage against the oracle's stage applied to the same input, then
    the chain bit reversal + stages 0 .. log2(n) - 1 against the FFT."""
    import torch
    rng = np.random.default_rng(n + 2)
    rows = 3
    z = _complex(rng, rows, n)
    a = _dev(dev, z, np.complex64)
    b = torch.empty_like(a)
    _call(ctr, "ctr_fft_bit_reverse", ctr.ptr(a), n, rows)
    for s in range(n.bit_length() - 1):
        read = _host(a)
        _call(ctr, "ctr_fft_single_stage", ctr.ptr(a), ctr.ptr(b), n, rows, s)
        got = _host(b)
        ref = np.stack([orc.fft_single_stage(r.astype(np.complex128), s) for r in read])
        assert rel_max(got, ref) < 4e-7, (n, s)
        a, b = b, a
    fft = np.stack([orc.fft_radix2(r.astype(np.complex128)) for r in z])
    assert rel_max(_host(a), fft) < 2e-6 * max(1, np.log2(n))
    # the same chain as one ctr_fft_radix2 launch
    c = _dev(dev, z, np.complex64)
    _call(ctr, "ctr_fft_radix2", ctr.ptr(c), n, rows)
    assert rel_max(_host(c), fft) < 2e-6 * max(1, np.log2(n))


def test_single_stage_rejects_aliasing_and_bad_steps(ctr, dev):
    import torch
    t = torch.zeros((2, 16), dtype=torch.complex64, device=dev)
    lib = ctr.lib()
    assert lib.ctr_fft_single_stage(None, ctr.ptr(t), ctr.ptr(t), 16, 2, 0) == -1
    u = torch.zeros_like(t)
    assert lib.ctr_fft_single_stage(None, ctr.ptr(t), ctr.ptr(u), 16, 2, 4) == -1
    assert lib.ctr_fft_single_stage(None, ctr.ptr(t), ctr.ptr(u), 12, 2, 0) == -2


# --- FBP2D::apply_filter (FBP2D.cpp:124-144) ---------------------------------

@pytest.mark.parametrize("n", [8, 100, 512, 2048])
@pytest.mark.parametrize("ftype", [0, 1, 2])
def test_apply_filter_vs_oracle_gain(ctr, orc, dev, n, ftype):
    """Shifted spectrum x H[i], H[i] = n g((float)i/n - 0.5) (FP64 gain of
    the oracle, FastFourierTransformImage.cpp:35-37,45)."""
    rng = np.random.default_rng(n * 3 + ftype)
    z = _complex(rng, 5, n)
    t = _dev(dev, z, np.complex64)
    _call(ctr, "ctr_apply_filter", ctr.ptr(t), n, 5, ftype)
    ref = z.astype(np.complex128) * orc.filter_gain(n, ftype)[None, :]
    got = _host(t)
    assert np.abs(got - ref).max() <= 3e-7 * np.abs(ref).max()
    assert np.all(got[:, n // 2] == 0)  # DC of the shifted spectrum


def test_apply_filter_rejects_unknown_filter(ctr, dev):
    import torch
    t = torch.zeros((1, 8), dtype=torch.complex64, device=dev)
    with pytest.raises(ctr.CtrError, match="unknown filter"):
        _call(ctr, "ctr_apply_filter", ctr.ptr(t), 8, 1, 3)


# --- detector / volume passes (FBP3D.cpp:333-387,444-458) --------------------

@pytest.mark.parametrize("layout", [0, 1])
def test_fdk_weight_standalone_vs_oracle(ctr, orc, dev, layout):
    rng = np.random.default_rng(30 + layout)
    N, nv, nu = 7, 24, 40
    sod, odd, w = 300.0, 200.0, 180.0
    p = rng.standard_normal((N, nv, nu)).astype(np.float32)
    ref = orc.fdk_weight(p, sod, odd, w)
    src = p if layout == 0 else np.ascontiguousarray(p.transpose(1, 0, 2))
    t = _dev(dev, src)
    _call(ctr, "ctr_fdk_weight", ctr.ptr(t), layout, nu, nv, N, ctypes.c_double(sod),
          ctypes.c_double(odd), ctypes.c_double(w))
    got = _host(t)
    if layout == 1:
        got = got.transpose(1, 0, 2)
    # FP64 weight times the FP32 sample, rounded once to FP32
    assert np.abs(got - ref).max() <= 2 ** -24 * np.abs(ref).max()


def test_log_normalize_vs_oracle(ctr, orc, dev):
    rng = np.random.default_rng(31)
    base, floor = 1e5, 1.0
    p = rng.uniform(0.0, 6.0, size=5000)
    I = (base * np.exp(-p)).astype(np.float32)
    I[:50] = 0.0          # dark pixels: clamped to the floor
    I[50:60] = -3.0
    t = _dev(dev, I)
    _call(ctr, "ctr_log_normalize", ctr.ptr(t), ctypes.c_size_t(I.size), ctypes.c_float(base),
          ctypes.c_float(floor))
    got = _host(t)
    ref = orc.log_normalize(I.astype(np.float64), base, floor)
    assert np.abs(got - ref).max() < 2e-6
    assert np.allclose(got[:60], np.log(base), rtol=1e-6)


def test_clip_and_fill(ctr, dev):
    rng = np.random.default_rng(32)
    v = rng.standard_normal(12345).astype(np.float32)
    v[7] = -0.0
    t = _dev(dev, v)
    _call(ctr, "ctr_clip_negative", ctr.ptr(t), ctypes.c_size_t(v.size))
    got = _host(t)
    assert np.array_equal(got, np.maximum(v, np.float32(0)))
    assert got.min() >= 0.0
    _call(ctr, "ctr_fill", ctr.ptr(t), ctypes.c_size_t(v.size), ctypes.c_float(-2.5))
    assert np.all(_host(t) == np.float32(-2.5))


# --- FBP3D::render_ellipsoid / generate_shepplogan (FBP3D.cpp:58-72,609-628) --

def test_render_ellipsoids_vs_oracle(ctr, orc, dev):
    """Seeded random ellipsoids (rotated, overlapping, some leaving the
    volume) voxelised on the GPU against the oracle voxeliser.  Voxels whose
    centre lies within rounding of an ellipsoid surface may differ (the GPU
    multiplies by 1/semi-axis, the oracle divides); any such voxel differs by
    exactly one ellipsoid's density."""
    import torch
    X, Y, Z = 61, 47, 53
    ell = ctr.random_ellipsoids(5, 25, (X, Y, Z), 1.0, (2, 30), inner=1.1)
    vol = torch.zeros((Z, Y, X), dtype=torch.float32, device=dev)
    ctr.render_ellipsoids(ell, vol, float(X))
    got = _host(vol).astype(np.float64)
    ref = orc.render_ellipsoids(ell, (X, Y, Z), float(X))
    diff = np.abs(got - ref) > 1e-5
    assert ref.max() > 0 and np.count_nonzero(ref) > 1000
    assert diff.sum() <= 1e-4 * diff.size
    dens = ell[:, 9]
    for d in np.abs(got - ref)[diff]:
        assert np.min(np.abs(dens - d)) < 1e-5
    # accumulate: rendering twice adds the phantom again
    ctr.render_ellipsoids(ell, vol, float(X))
    assert np.allclose(_host(vol), 2 * got.astype(np.float32), rtol=1e-6, atol=1e-6)
