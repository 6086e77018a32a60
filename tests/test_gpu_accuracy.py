"""Accuracy of each workload's reconstruction against the FP64 oracle pipeline
on the same inputs: relative RMSE, max-abs error, max-rel error (over
max |ref|) and PSNR = 20 log10(max |ref| / RMSE), for the FP32 form and the
FP16 / BF16 storage forms (filtered projections stored as 16-bit
value-and-step texels with a per-view scale; FP32 interpolation and
accumulation).  Errors are taken before the negative clip.  The oracle pipeline is the FP64 restatement:
FDK weight, reference radix-2 ramp filter and voxel loop in FP64, with FP32
arrays between the stages (the same interchange format the product uses).

* par2d_256: 2D parallel, 256^2 from 180 views of 367 bins; analytic
  sinogram of 10 seeded ellipses; Ram-Lak, FFT 512; the whole image.
* fan2d_512: 2D fan, 512^2 from 720 views of 768 bins (SOD 500, SDD 1000);
  20 seeded ellipses through the ray-driven forward projector; textbook
  Shepp-Logan window (the reference's COSINE filter), FFT 2048; whole image.
* fdk_512: 512^3 from 720 views of 768^2, the benchmark input (30 seeded
  ellipsoids, ray-driven projections, Poisson noise at I0 = 1e5, log);
  FDK weight + Ram-Lak (FFT 2048) + backprojection; two 32^3 crops inside
  the phantom (the centre, and one 195 voxels off the axis near the top of
  the volume, at a high cone angle).
The 1440^3 workload's three forms are checked in
test_gpu_fullsize.test_1440_storage_forms_vs_oracle.

Each case prints one '[accuracy]' line per form.  Gates: FP32 max-rel
<= 1e-4; 16-bit forms rel RMSE <= 1e-3.
"""
import numpy as np
import pytest

from conftest import error_metrics as metrics, format_metrics

pytestmark = pytest.mark.gpu


def report(case, name, m):
    print(f"[accuracy] {case} {name}: {format_metrics(m)}", flush=True)


def check(case, name, m):
    report(case, name, m)
    if name == "fp32":
        assert m["max_rel"] <= 1e-4, (case, name, m)
    else:
        assert m["rel_rmse"] <= 1e-3, (case, name, m)


def forms(ctr):
    return (("fp32", ctr.F32), ("fp16", ctr.F16), ("bf16", ctr.BF16))


def _t(dev, a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).to(dev)


def test_accuracy_par2d_256(ctr, orc, dev):
    g = ctr.config_geometry("par2d_256")
    ell = ctr.random_ellipsoids(101, 10, (g.X, 1, g.Z), 1.0, (5, 60), planar=True)
    sino = orc.analytic_parallel(ell, g.detector_width, g.nu, 1, g.n_views,
                                 g.rotation_count, 0.0).astype(np.float32)
    nfft, scale = 512, g.fbp_scale(512)
    q = orc.filter_rows(sino.reshape(-1, g.nu), orc.RAM_LAK, nfft).reshape(sino.shape)
    ref = orc.backproject_parallel(q, g.detector_width, g.volume_width, g.rotation_count,
                                   0.0, scale, (g.X, 1, g.Z))
    for name, dt in forms(ctr):
        staged, vs = ctr.filter_projections(_t(dev, sino), g, ctr.RAM_LAK, nfft, dt, fdk=False)
        img = ctr.backproject(staged, g, dt, view_scale=vs, scale=scale).cpu().numpy()
        check("par2d_256", name, metrics(img, ref))


def test_accuracy_fan2d_512(ctr, orc, dev):
    g = ctr.config_geometry("fan2d_512")
    ell = ctr.random_ellipsoids(202, 20, (g.X, 1, g.Z), g.voxel, (5, 60), planar=True)
    img = orc.render_ellipsoids(ell, (g.X, 1, g.Z), g.volume_width).astype(np.float32)
    sino = ctr.forward_project(_t(dev, img), g, step=1.0)
    sino_h = sino.cpu().numpy()
    nfft, scale = 2048, g.fbp_scale(2048)
    w = orc.fdk_weight(sino_h, g.source_distance, g.detector_distance, g.detector_width)
    q = orc.filter_rows(w.reshape(-1, g.nu), orc.COSINE, nfft).reshape(w.shape)
    ref = orc.backproject_cone(q, g.source_distance, g.detector_distance, g.detector_width,
                               g.volume_width, 1.0, 0.0, scale, (g.X, 1, g.Z))
    for name, dt in forms(ctr):
        staged, vs = ctr.filter_projections(sino, g, ctr.COSINE, nfft, dt, fdk=True)
        rec = ctr.backproject(staged, g, dt, view_scale=vs, scale=scale).cpu().numpy()
        check("fan2d_512", name, metrics(rec, ref))


def test_accuracy_fdk_512(ctr, orc, dev):
    import torch
    g = ctr.config_geometry("fdk_512")
    ell = ctr.random_ellipsoids(20240602, 30, (g.X, g.Y, g.Z), g.voxel, (5, 120))
    ell[:, 9] *= 0.02  # density 1.0 = 0.02 / mm
    vol = torch.zeros((g.Z, g.Y, g.X), dtype=torch.float32, device=dev)
    ctr.render_ellipsoids(ell, vol, g.volume_width)
    proj = ctr.forward_project(vol, g, step=1.0)
    ctr.poisson_log(proj, 1e5, 20240602)
    del vol
    torch.cuda.synchronize()
    print("[accuracy] fdk_512 projections ready", flush=True)
    nfft, scale = 2048, g.fbp_scale(2048)
    boxes = (((240, 272), (240, 272), (240, 272)),   # centre
             ((80, 112), (440, 472), (128, 160)))    # off axis, high cone angle
    got = {}
    for name, dt in forms(ctr):
        staged, vs = ctr.filter_projections(proj, g, ctr.RAM_LAK, nfft, dt, fdk=True)
        rec = ctr.backproject(staged, g, dt, view_scale=vs, scale=scale)
        got[name] = [rec[k0:k1, j0:j1, i0:i1].cpu().numpy()
                     for (i0, i1), (j0, j1), (k0, k1) in boxes]
        del staged, rec
        torch.cuda.synchronize()
    ph = proj.cpu().numpy()
    del proj
    w = orc.fdk_weight(ph, g.source_distance, g.detector_distance, g.detector_width)
    del ph
    q = orc.filter_rows(w.reshape(-1, g.nu), orc.RAM_LAK, nfft).reshape(w.shape)
    del w
    print("[accuracy] fdk_512 oracle filter done", flush=True)
    refs = [orc.backproject_cone_region(q, g.source_distance, g.detector_distance,
                                        g.detector_width, g.volume_width, 1.0, 0.0, scale,
                                        (g.X, g.Y, g.Z), box) for box in boxes]
    for name, _ in forms(ctr):
        for tag, a, r in zip(("centre", "off-axis"), got[name], refs):
            check(f"fdk_512 {tag} 32^3", name, metrics(a, r))
