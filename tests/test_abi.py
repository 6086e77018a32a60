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
    assert lib.ctr_forward_project(None, None, 0, fake, 0, 8, 8, 100.0, 1.0, 10.0, 10.0,
                                   1.0, 0.0, fake, 0, 8, 8, 4, 1.0) == -1
    assert lib.ctr_render_ellipsoids(None, None, fake, 5000, fake, 4, 4, 4, 4.0) == -1
    assert lib.ctr_staged_layout(0, 4, 4, 0, None, None, None) == -1
    # staged formats are CTR_F32, CTR_F16 and CTR_BF16 only
    for dt in (3, 4, -1):
        assert lib.ctr_staged_layout(8, 8, 4, dt, None, None, None) == -1
        assert "unknown staged dtype" in ctr.last_error()
        assert lib.ctr_backproject(None, None, 0, fake, dt, fake, 8, 8, 4, 100.0, 1.0, 10.0,
                                   10.0, 1.0, 0.0, fake, 8, 8, 8, 1.0, 0, 0) == -1
        assert "unknown staged dtype" in ctr.last_error()


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


def test_workspace_options_without_device(ctr):
    """The default workspace (NULL) reports the load-time options and refuses
    changes; unknown options and values are rejected; no device is touched."""
    lib = ctr.lib()
    assert lib.ctr_workspace_get_option(None, ctr.OPTION_BACKPROJECT_KERNEL) in (
        ctr.KERNEL_AUTO, ctr.KERNEL_LDS, ctr.KERNEL_GATHER)
    assert lib.ctr_workspace_set_option(None, ctr.OPTION_BACKPROJECT_KERNEL,
                                        ctr.KERNEL_GATHER) == -1
    assert "default workspace" in ctr.last_error()
    for bad in (3, 9, -1):
        assert lib.ctr_workspace_get_option(None, bad) == -1
        assert "unknown option" in ctr.last_error()


@pytest.mark.parametrize("env,option,value", [
    ({"CTR_BACKPROJECT_KERNEL": "lds"}, 0, 1), ({"CTR_BACKPROJECT_KERNEL": "gather"}, 0, 2),
    ({"CTR_BACKPROJECT_KERNEL": ""}, 0, 0),
    ({"CTR_FORWARD_KERNEL": "ray"}, 1, 3), ({"CTR_FORWARD_ADDRESSING": "pointer"}, 2, 4),
    ({}, 0, 0), ({}, 1, 0), ({}, 2, 0)])
def test_options_read_from_environment_at_load(ctr, env, option, value):
    """The environment is read once, when the library is loaded: changing it
    afterwards changes nothing."""
    code = ("import ctypes, os, sys; L = ctypes.CDLL(sys.argv[1]); "
            "a = L.ctr_workspace_get_option(None, int(sys.argv[2])); "
            "os.environ['CTR_BACKPROJECT_KERNEL'] = 'gather'; "
            "os.environ['CTR_FORWARD_KERNEL'] = 'ray'; "
            "os.environ['CTR_FORWARD_ADDRESSING'] = 'pointer'; "
            "b = L.ctr_workspace_get_option(None, int(sys.argv[2])); print(a, b)")
    base = {k: v for k, v in os.environ.items() if not k.startswith("CTR_")}
    out = subprocess.run(["python", "-c", code, ctr.LIB_PATH, str(option)],
                         env=dict(base, **env), capture_output=True, text=True,
                         check=True).stdout.split()
    assert [int(x) for x in out] == [value, value]


def test_size_limits_without_device(ctr):
    """Degenerate and over-limit sizes are rejected on the host, before any
    device call: empty view sets, sizes of 2^20 and more, a staged view past
    the backprojector's 32-bit in-view offsets (2^31 bytes), packed counts
    outside 1..16 and a non-positive forward step."""
    lib = ctr.lib()
    fake = ctypes.c_void_p(64)
    bp = lambda nu, nv, n, X, dt=0, vs=None: lib.ctr_backproject(
        None, None, 0, fake, dt, vs, nu, nv, n, 500.0, 500.0, 600.0, 200.0, 1.0, 0.0,
        fake, X, X, X, 1.0, 0, 0)
    assert bp(64, 64, 0, 32) == -1
    assert "sizes must be positive" in ctr.last_error()
    assert bp(64, 64, 8, 1 << 20) == -1
    assert "below 2^20" in ctr.last_error()
    # 24000^2 detector: a staged view of 24032 x 24003 texels, 2.3e9 bytes
    assert bp(24000, 24000, 8, 32) == -1
    assert "32-bit in-view offsets" in ctr.last_error()
    assert bp(24000, 24000, 8, 32, dt=1, vs=fake) == -1
    jobs = (ctr.BackprojectJob * 17)()
    assert lib.ctr_backproject_packed(None, None, 0, jobs, 0) == -1
    assert lib.ctr_backproject_packed(None, None, 0, jobs, 17) == -1
    assert lib.ctr_backproject_packed(None, None, 0, None, 1) == -5
    assert lib.ctr_backproject_packed(None, None, 7, jobs, 1) == -1
    assert lib.ctr_forward_project(None, None, 0, fake, 8, 8, 8, 100.0, 1.0, 10.0, 10.0,
                                   1.0, 0.0, fake, 0, 8, 8, 4, 0.0) == -1
    assert "step must be positive" in ctr.last_error()
