"""The reference's 2D programs on the drop-in FBP2D / FFFT classes, stage by
stage against the oracle.

examples/fbp_parallel2d_main.cpp runs the call sequence of
Mains/fbp_parallel2d_main.cpp:29-41 and examples/fbp_fan2d_equidistance_main.cpp
that of Mains/fbp_fan2d_equidistance_main.cpp:32-49 (forward projection,
fan -> parallel rebinning, create_complex_texture_from_real -> fft_radix2 ->
fft_shift -> FBP2D::apply_filter -> inverse_fft_shift -> inverse_fft_radix2 ->
blit_texture_complex_to_real_r -> FBP2D::project_backward).  The reference
reads a PNG (shepplogan.png / lenna.png) through GraphicsCortex, which is
absent; the input here is a seeded ellipse phantom of the same size.

Each stage's GPU output is compared with the oracle applied to the GPU's own
previous stage (so each comparison isolates one stage), and the whole chain
with the oracle chain from the image.  FP32 tolerance 1e-4 max relative
error (1e-5 for the FFT chain, which is a few FP32 roundings deep).
"""
import math
import os
import subprocess

import numpy as np
import pytest

from conftest import rel_max

pytestmark = pytest.mark.gpu

SOD, ODD, DET, VOLW = (float(np.float32(x)) for x in (730.87, 669.04, 409.60, 213.84))


def _exe(ctr, name):
    exe = os.path.join(os.path.dirname(ctr.LIB_PATH), "examples", name)
    assert os.path.exists(exe), "examples are built by __graft_entry__.build()"
    return exe


def _image(ctr, orc, W, seed):
    ell = ctr.random_ellipsoids(seed, 10, (W, 1, W), 1.0, (W / 40, W / 5), planar=True)
    return orc.render_ellipsoids(ell, (W, 1, W), float(W))[:, 0, :].astype(np.float32)


def _run(exe, img, W, tmp_path, filt):
    src = tmp_path / "image.raw"
    img.tofile(src)
    r = subprocess.run([exe, str(src), str(W), str(tmp_path), str(filt)], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    return lambda name, ch=1: np.fromfile(tmp_path / f"{name}.raw", dtype=np.float32).reshape(
        (W, W, ch) if ch > 1 else (W, W))


def _oracle_filter_chain(orc, sino, ftype):
    """create_complex_texture_from_real -> fft_radix2 -> fft_shift -> x H ->
    inverse_fft_shift -> inverse_fft_radix2 -> real part, per row, FP64
    (no zero padding: the rows are W long, a power of two)."""
    n = sino.shape[1]
    gain = orc.filter_gain(n, ftype)
    out = np.empty(sino.shape)
    for r, row in enumerate(sino.astype(np.float64)):
        z = orc.fft_shift(orc.fft_radix2(row + 0j), n // 2) * gain
        out[r] = orc.inverse_fft_radix2(orc.fft_shift(z, -(n // 2))).real
    return out


def _oracle_backward(orc, sino, W):
    """FBP2D::project_backward (FBP2D.cpp:96-122): one full turn with a
    0.006 rad offset, plane width W sqrt(2), sum / N."""
    N = sino.shape[0]
    return orc.backproject_parallel(sino.astype(np.float32)[:, None, :], W * math.sqrt(2),
                                    float(W), 1.0, 0.006, 1.0 / N, (W, 1, W))[:, 0, :]


@pytest.mark.parametrize("W,filt", [(256, 0), (256, 2), (1024, 0)])
def test_parallel2d_main_vs_oracle(ctr, orc, dev, tmp_path, W, filt):
    img = _image(ctr, orc, W, 11 + W)
    out = _run(_exe(ctr, "fbp_parallel2d_main"), img, W, tmp_path, filt)
    sino, filtered, rec = out("sinogram"), out("filtered"), out("reconstruction")
    # FBP2D::project_forward: parallel rays, plane W sqrt(2), W views, 1 turn
    ref_sino = orc.forward_parallel(img[:, None, :], float(W), W * math.sqrt(2), W, 1, W,
                                    1.0, 0.0, 1.0)[:, 0, :]
    assert rel_max(sino, ref_sino) < 1e-4
    assert rel_max(filtered, _oracle_filter_chain(orc, sino, filt)) < 1e-5
    assert rel_max(rec, _oracle_backward(orc, filtered, W)) < 1e-4
    # whole chain from the image
    full = _oracle_backward(orc, _oracle_filter_chain(orc, ref_sino, filt), W)
    assert rel_max(rec, full) < 1e-4
    rgba = out("reconstruction_rgba", 4)
    assert all(np.array_equal(rgba[..., c], rec) for c in range(4))
    print(f"[parallel2d {W}] recon max {rec.max():.3f}, phantom max {img.max():.3f}")


@pytest.mark.parametrize("W", [256, 512])
def test_fan2d_equidistance_main_vs_oracle(ctr, orc, dev, tmp_path, W):
    img = _image(ctr, orc, W, 29 + W)
    out = _run(_exe(ctr, "fbp_fan2d_equidistance_main"), img, W, tmp_path, 0)
    par = out("sinogram_parallel")
    ref_par = orc.forward_parallel(img[:, None, :], float(W), W * math.sqrt(2), W, 1, W,
                                   1.0, 0.0, 1.0)[:, 0, :]
    assert rel_max(par, ref_par) < 1e-4
    # FBP2D::project_forward_equiangular: flat-detector fan, the main's
    # geometry in mm over a field VOLW wide
    fan = out("sinogram_fan")
    ref_fan = orc.forward_cone(img[:, None, :], VOLW, SOD, ODD, DET, W, 1, W, 1.0, 0.0,
                               1.0)[:, 0, :]
    assert rel_max(fan, ref_fan) < 1e-4
    rebinned = out("sinogram_rebinned")
    ref_reb = orc.rebin_fan_to_parallel(fan, SOD, ODD, DET, 1.0, W, VOLW * math.sqrt(2))
    assert rel_max(rebinned, ref_reb) < 1e-5
    filtered = out("filtered")
    assert rel_max(filtered, _oracle_filter_chain(orc, rebinned, 0)) < 1e-5
    rec = out("reconstruction")
    assert rel_max(rec, _oracle_backward(orc, filtered, W)) < 1e-4
    ref_full = _oracle_backward(orc, _oracle_filter_chain(
        orc, orc.rebin_fan_to_parallel(ref_fan.astype(np.float32), SOD, ODD, DET, 1.0, W,
                                       VOLW * math.sqrt(2)), 0), W)
    assert rel_max(rec, ref_full) < 1e-4
