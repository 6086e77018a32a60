"""CPU-side checks of the C ABI: the library loads, exports every entry point
include/ctrecon.h declares, and rejects bad arguments before touching a
device (no GPU is used here)."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "ctrecon.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ctr_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_expected_surface():
    syms = declared_symbols()
    for name in ("ctr_fft_radix2", "ctr_inverse_fft_radix2", "ctr_fft_shift",
                 "ctr_apply_filter", "ctr_filter_rows", "ctr_filter_projections",
                 "ctr_backproject", "ctr_forward_project", "ctr_fdk_weight",
                 "ctr_clip_negative", "ctr_last_error"):
        assert name in syms


def test_library_exports_every_declared_symbol(ctr):
    lib = ctr.lib()
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", ctr.LIB_PATH],
                         capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r" T (ctr_[a-z0-9_]+)", out))
    assert set(declared_symbols()) <= exported
    # the Python binding covers the whole ABI
    assert set(declared_symbols()) == set(ctr.SIGNATURES)


def test_argument_errors_without_device(ctr):
    lib = ctr.lib()
    fake = ctypes.c_void_p(64)
    assert lib.ctr_fft_radix2(None, None, 8, 1) == -5
    assert lib.ctr_fft_radix2(None, fake, 12, 1) == -2
    assert "not a power of 2" in ctr.last_error()
    assert lib.ctr_inverse_fft_radix2(None, fake, 6, 1) == -2
    assert lib.ctr_apply_filter(None, fake, 8, 1, 3) == -1
    assert "unknown filter" in ctr.last_error()
    assert lib.ctr_filter_rows(None, fake, fake, 100, 2, 96, 0) == -2
    assert lib.ctr_filter_rows(None, fake, fake, 100, 2, 64, 0) == -1
    assert lib.ctr_filter_rows(None, fake, fake, 100, 2, 16384, 0) == -1
    assert lib.ctr_fft_single_stage(None, fake, fake, 8, 1, 0) == -1  # read == write
    # cone source inside the volume
    assert lib.ctr_backproject(None, None, 0, fake, 0, None, 8, 8, 4, 1.0, 1.0, 10.0,
                               10.0, 1.0, 0.0, fake, 8, 8, 8, 1.0, 0, 0) == -1
    assert "source distance" in ctr.last_error()
    assert lib.ctr_backproject(None, None, 0, fake, 1, None, 8, 8, 4, 100.0, 1.0, 10.0,
                               10.0, 1.0, 0.0, fake, 8, 8, 8, 1.0, 0, 0) == -5
    assert lib.ctr_backproject(None, None, 5, fake, 0, None, 8, 8, 4, 100.0, 1.0, 10.0,
                               10.0, 1.0, 0.0, fake, 8, 8, 8, 1.0, 0, 0) == -1
    assert lib.ctr_forward_project(None, 0, fake, 0, 8, 8, 100.0, 1.0, 10.0, 10.0, 1.0,
                                   0.0, fake, 0, 8, 8, 4, 1.0) == -1
    assert lib.ctr_render_ellipsoids(None, None, fake, 5000, fake, 4, 4, 4, 4.0) == -1
    assert lib.ctr_staged_layout(0, 4, 4, 0, None, None, None) == -1


def test_staged_layout_sizes(ctr):
    pu, pv, nbytes = ctr.staged_layout(768, 768, 720, ctr.F32)
    assert pu >= 768 + 3 and pu % 4 == 0 and pv == 768 + 3
    assert nbytes == pu * pv * 720 * 4
    assert ctr.staged_layout(768, 768, 720, ctr.F16) == (pu, pv, nbytes)


def test_product_path_fails_loudly_without_library(ctr, monkeypatch):
    monkeypatch.setattr(ctr, "_LIB", None)
    monkeypatch.setattr(ctr, "LIB_PATH", os.path.join(ROOT, "no_such_libctrecon.so"))
    with pytest.raises(RuntimeError, match="no CPU fallback"):
        ctr.lib()


def test_interpolation_mode_switch(ctr):
    """ctr_set_interpolation: returns the previous mode, rejects unknown ones
    and leaves the mode unchanged then; no device is touched."""
    lib = ctr.lib()
    start = lib.ctr_get_interpolation()
    assert start in (ctr.INTERP_FAST, ctr.INTERP_FP32)
    try:
        assert ctr.set_interpolation(ctr.INTERP_FP32) == start
        assert ctr.get_interpolation() == ctr.INTERP_FP32
        assert lib.ctr_set_interpolation(7) == -1
        assert "unknown mode" in ctr.last_error()
        assert ctr.get_interpolation() == ctr.INTERP_FP32
        with pytest.raises(ctr.CtrError):
            ctr.set_interpolation(-1)
        assert ctr.set_interpolation(ctr.INTERP_FAST) == ctr.INTERP_FP32
        assert ctr.get_interpolation() == ctr.INTERP_FAST
    finally:
        ctr.set_interpolation(start)


@pytest.mark.parametrize("value,mode", [("fp32", 1), ("fast", 0), ("", 1)])
def test_interpolation_mode_from_environment(ctr, value, mode):
    code = ("import ctypes, sys; L = ctypes.CDLL(sys.argv[1]); "
            "print(L.ctr_get_interpolation())")
    env = dict(os.environ, CTR_INTERPOLATION=value)
    out = subprocess.run(["python", "-c", code, ctr.LIB_PATH], env=env,
                         capture_output=True, text=True, check=True).stdout
    assert int(out.strip()) == mode


def test_size_limits_without_device(ctr):
    """Degenerate and over-limit sizes are rejected on the host, before any
    device call: empty view sets, sizes of 2^20 and more, a staged view past
    the backprojector's 2^22-texel offsets, packed counts outside 1..16 and a
    non-positive forward step."""
    lib = ctr.lib()
    fake = ctypes.c_void_p(64)
    bp = lambda nu, nv, n, X, dt=0, vs=None: lib.ctr_backproject(
        None, None, 0, fake, dt, vs, nu, nv, n, 500.0, 500.0, 600.0, 200.0, 1.0, 0.0,
        fake, X, X, X, 1.0, 0, 0)
    assert bp(64, 64, 0, 32) == -1
    assert "sizes must be positive" in ctr.last_error()
    assert bp(64, 64, 8, 1 << 20) == -1
    assert "below 2^20" in ctr.last_error()
    # 2100 x 2100 detector: a staged view of 2112 x 2103 texels > 2^22
    assert bp(2100, 2100, 8, 32) == -1
    assert "2^22" in ctr.last_error()
    assert bp(2100, 2100, 8, 32, dt=1, vs=fake) == -1
    jobs = (ctr.BackprojectJob * 17)()
    assert lib.ctr_backproject_packed(None, None, 0, jobs, 0) == -1
    assert lib.ctr_backproject_packed(None, None, 0, jobs, 17) == -1
    assert lib.ctr_backproject_packed(None, None, 0, None, 1) == -5
    assert lib.ctr_backproject_packed(None, None, 7, jobs, 1) == -1
    assert lib.ctr_forward_project(None, 0, fake, 8, 8, 8, 100.0, 1.0, 10.0, 10.0, 1.0,
                                   0.0, fake, 0, 8, 8, 4, 0.0) == -1
    assert "step must be positive" in ctr.last_error()
