"""2D paths on the GPU vs the FP64 oracle: FBP2D's reference-convention
backprojection, the 2D parallel FBP workload (256^2, 180 views, 367 bins),
the 2D fan workload (512^2, 720 views, 768 bins, SOD 500 / SDD 1000) and the
fan -> parallel rebinning.  FP32 tolerance: 1e-4 max relative error."""
import math

import numpy as np
import pytest

from conftest import rel_max

pytestmark = pytest.mark.gpu


def _t(dev, a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).to(dev)


def test_fbp2d_reference_backprojection(ctr, orc, dev):
    """FBP2D::project_backward (FBP2D.cpp:96-122): one full turn, +0.006 rad,
    plane width W*sqrt(2), sum / N."""
    W, N, nu = 96, 120, 136
    rng = np.random.default_rng(1)
    sino = rng.standard_normal((N, nu)).astype(np.float32)
    g = ctr.Geometry(ctr.PARALLEL, nu=nu, nv=1, n_views=N, X=W, Y=1, Z=W,
                     detector_width=W * math.sqrt(2), volume_width=float(W),
                     rotation_count=1.0, rotation_offset=0.006)
    staged, _ = ctr.stage_projections(_t(dev, sino[:, None, :]), g)
    got = ctr.backproject(staged, g, scale=1.0 / N).cpu().numpy()[:, 0, :]
    ref = orc.backproject_parallel(sino[:, None, :], W * math.sqrt(2), float(W), 1.0, 0.006,
                                   1.0 / N, (W, 1, W))[:, 0, :]
    assert rel_max(got, ref) < 1e-4


def test_parallel_2d_workload(ctr, orc, dev):
    g = ctr.config_geometry("par2d_256")
    ell = ctr.random_ellipsoids(101, 10, (g.X, 1, g.Z), 1.0, (5, 60), planar=True)
    sino = orc.analytic_parallel(ell, g.detector_width, g.nu, 1, g.n_views,
                                 g.rotation_count, 0.0).astype(np.float32)
    nfft = 512
    staged, _ = ctr.filter_projections(_t(dev, sino), g, ctr.RAM_LAK, nfft, fdk=False)
    img = ctr.backproject(staged, g, scale=g.fbp_scale(nfft)).cpu().numpy()
    q = orc.filter_rows(sino.reshape(-1, g.nu), 0, nfft).reshape(sino.shape)
    ref = orc.backproject_parallel(q.astype(np.float32), g.detector_width, g.volume_width,
                                   g.rotation_count, 0.0, g.fbp_scale(nfft), (g.X, 1, g.Z))
    assert rel_max(img, ref) < 1e-4
    truth = orc.render_ellipsoids(ell, (g.X, 1, g.Z), g.volume_width)
    mse = float(((img - truth) ** 2).mean())
    psnr = 20 * math.log10(truth.max() / math.sqrt(mse))
    print(f"[2d parallel] PSNR vs phantom {psnr:.1f} dB")
    assert psnr > 20.0


def test_fan_2d_workload(ctr, orc, dev):
    g = ctr.config_geometry("fan2d_512")
    ell = ctr.random_ellipsoids(202, 20, (g.X, 1, g.Z), g.voxel, (5, 60), planar=True)
    img = orc.render_ellipsoids(ell, (g.X, 1, g.Z), g.volume_width).astype(np.float32)
    sino = ctr.forward_project(_t(dev, img), g, step=1.0)
    ref_sino = orc.forward_cone(img, g.volume_width, g.source_distance, g.detector_distance,
                                g.detector_width, g.nu, 1, g.n_views, 1.0, 0.0, 1.0)
    sino_h = sino.cpu().numpy()
    assert rel_max(sino_h, ref_sino) < 1e-4
    nfft = 2048
    staged, _ = ctr.filter_projections(sino, g, ctr.COSINE, nfft, fdk=True)
    rec = ctr.backproject(staged, g, scale=g.fbp_scale(nfft)).cpu().numpy()
    w = orc.fdk_weight(sino_h, g.source_distance, g.detector_distance, g.detector_width)
    q = orc.filter_rows(w.reshape(-1, g.nu).astype(np.float32), orc.COSINE, nfft)
    ref = orc.backproject_cone(q.reshape(w.shape).astype(np.float32), g.source_distance,
                               g.detector_distance, g.detector_width, g.volume_width, 1.0, 0.0,
                               g.fbp_scale(nfft), (g.X, 1, g.Z))
    assert rel_max(rec, ref) < 1e-4
    mse = float(((rec - img) ** 2).mean())
    psnr = 20 * math.log10(img.max() / math.sqrt(mse))
    print(f"[2d fan] PSNR vs phantom {psnr:.1f} dB")
    assert psnr > 20.0


def test_rebin_fan_to_parallel(ctr, orc, dev):
    g = ctr.Geometry(ctr.CONE, nu=256, nv=1, n_views=360, X=1, Y=1, Z=1,
                     detector_width=300.0, volume_width=1.0, source_distance=400.0,
                     detector_distance=300.0)
    rng = np.random.default_rng(3)
    fan = rng.standard_normal((g.n_views, g.nu)).astype(np.float32)
    got = ctr.rebin_fan_to_parallel(_t(dev, fan), g, 300, 180.0).cpu().numpy()
    ref = orc.rebin_fan_to_parallel(fan, 400.0, 300.0, 300.0, 1.0, 300, 180.0)
    assert rel_max(got, ref) < 1e-5
