"""Backprojection at the benchmark sizes, checked against the FP64 oracle on
voxel crops (centre, corner, off-axis slab) and against size-independent
closed forms.

* FDK 512^3 from 720 views of 768^2 (SOD 500, SDD 1000, 0.5 mm voxels)
* FDK 1440 x 1470 x 1440 from 1440 views of 1536^2 (0.25 mm voxels)
Tolerances as in test_gpu_parity: FP32 <= 1e-4 max relative error on each
crop, FP16 storage <= 1e-3 relative RMSE.
"""
import math

import numpy as np
import pytest

from conftest import error_metrics, format_metrics, rel_max, rel_rmse

pytestmark = pytest.mark.gpu

CROPS_512 = [((248, 264), (248, 264), (248, 264)),   # centre
             ((0, 12), (500, 512), (0, 12)),          # corner, outside the FOV
             ((300, 312), (40, 52), (180, 192))]       # off axis, high cone angle
CROPS_1440 = [((716, 724), (731, 739), (716, 724)),
              ((100, 108), (1400, 1408), (1300, 1308))]


def _random_views(ctr, g, dev, seed):
    import torch
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    return torch.randn((g.n_views, g.nv, g.nu), generator=gen, device=dev,
                       dtype=torch.float32)


def _crop(vol, box):
    (i0, i1), (j0, j1), (k0, k1) = box
    return vol[k0:k1, j0:j1, i0:i1].cpu().numpy()


@pytest.mark.parametrize("workload,crops", [("fdk_512", CROPS_512), ("fdk_1440", CROPS_1440)])
def test_full_size_backprojection_crops(ctr, orc, dev, workload, crops):
    import torch
    g = ctr.config_geometry(workload)
    print(f"[fullsize] {workload}: {g.X}x{g.Y}x{g.Z}, {g.n_views} views", flush=True)
    q = _random_views(ctr, g, dev, 1234)
    staged, _ = ctr.stage_projections(q, g, ctr.F32)
    scale = g.fbp_scale(2048)
    vol = ctr.backproject(staged, g, ctr.F32, scale=scale)
    torch.cuda.synchronize()
    del staged
    qh = q.cpu().numpy()
    for box in crops:
        ref = orc.backproject_cone_region(qh, g.source_distance, g.detector_distance,
                                          g.detector_width, g.volume_width, 1.0, 0.0, scale,
                                          (g.X, g.Y, g.Z), box)
        got = _crop(vol, box)
        assert rel_max(got, ref) < 1e-4, box


def test_full_size_fp16_storage_crops(ctr, orc, dev):
    import torch
    g = ctr.config_geometry("fdk_512")
    q = _random_views(ctr, g, dev, 99)
    staged, vs = ctr.stage_projections(q, g, ctr.F16)
    vol = ctr.backproject(staged, g, ctr.F16, view_scale=vs, scale=1e-3)
    torch.cuda.synchronize()
    qh = q.cpu().numpy()
    peak = None
    for box in CROPS_512[:1] + CROPS_512[2:]:
        ref = orc.backproject_cone_region(qh, g.source_distance, g.detector_distance,
                                          g.detector_width, g.volume_width, 1.0, 0.0, 1e-3,
                                          (g.X, g.Y, g.Z), box)
        assert rel_rmse(_crop(vol, box), ref) < 1e-3
        peak = max(peak or 0.0, float(np.abs(ref).max()))
    # corner crop: its projections leave the detector sideways in some views
    # (the values there are small, so the bound is relative to the centre's
    # peak rather than to the crop itself)
    box = CROPS_512[1]
    ref = orc.backproject_cone_region(qh, g.source_distance, g.detector_distance,
                                      g.detector_width, g.volume_width, 1.0, 0.0, 1e-3,
                                      (g.X, g.Y, g.Z), box)
    assert np.abs(_crop(vol, box) - ref).max() < 1e-3 * peak


def test_1440_storage_forms_vs_oracle(ctr, orc, dev):
    """The 1440 x 1470 x 1440 scan (1440 views of 1536^2, FFT 4096): a body
    ellipsoid (semi-axes 140 / 170 / 140 mm, density 0.02 / mm) with 60 random
    ellipsoids inside the volume, forward projected, FDK-weighted and
    ramp-filtered with the filtered projections stored as FP32, FP16 and BF16
    (per-view scale, FP32 interpolation and accumulation, FP64 geometry).
    Each reconstruction is checked against the FP64 oracle backprojection of
    the FP32 filtered projections on the central 64^3 crop: FP32 max-rel
    <= 1e-4, 16-bit rel RMSE <= 1e-3.  (Measured about 8e-7 for FP32; 4.7e-6
    and 3.8e-5 for FP16 and BF16.  Without the body the crop is empty
    space, where the same absolute error over a reference 160 times smaller
    reads as 1e-3 and more.)"""
    import torch
    g = ctr.config_geometry("fdk_1440")
    ell = ctr.random_ellipsoids(3, 60, (g.X, g.Y, g.Z), g.voxel, (5, 240))
    ell[:, 9] *= 0.02
    body = np.zeros((1, 10))
    body[0, 3:6] = [140.0, 170.0, 140.0]
    body[0, 9] = 0.02
    phantom = torch.zeros((g.Z, g.Y, g.X), dtype=torch.float32, device=dev)
    ctr.render_ellipsoids(np.concatenate([body, ell]), phantom, g.volume_width)
    proj = ctr.forward_project(phantom, g, step=2.0)
    del phantom
    torch.cuda.synchronize()
    print("[1440^3] projections ready", flush=True)
    box = ((688, 752), (703, 767), (688, 752))
    scale = g.fbp_scale(4096)
    forms = (("fp32", ctr.F32), ("fp16", ctr.F16), ("bf16", ctr.BF16))
    got, filtered = {}, None
    ws = ctr.Workspace()
    for name, dt in forms:
        staged, vs = ctr.filter_projections(proj, g, ctr.RAM_LAK, 4096, dt, fdk=True, ws=ws)
        vol = ctr.backproject(staged, g, dt, view_scale=vs, scale=scale, ws=ws)
        got[name] = _crop(vol, box)
        if dt == ctr.F32:
            pu, pv, _ = ctr.staged_layout(g.nu, g.nv, g.n_views)
            st = staged.view(torch.float32).view(g.n_views, pv, pu)
            filtered = st[:, 1:g.nv + 1, 1:g.nu + 1].cpu().numpy()
        del staged, vol
        torch.cuda.synchronize()
    ws.close()
    del proj
    ref = orc.backproject_cone_region(filtered, g.source_distance, g.detector_distance,
                                      g.detector_width, g.volume_width, 1.0, 0.0, scale,
                                      (g.X, g.Y, g.Z), box)
    errs = {name: error_metrics(got[name], ref) for name in got}
    for name, _ in forms:
        print(f"[accuracy] fdk_1440 centre 64^3 {name}: {format_metrics(errs[name])}",
              flush=True)
    assert errs["fp32"]["max_rel"] < 1e-4
    for name, _ in forms[1:]:
        assert errs[name]["rel_rmse"] < 1e-3, name


def test_full_size_constant_projection_closed_form(ctr, dev):
    """q = 1 on the whole detector: every voxel whose projection stays inside
    the panel in all views receives scale * sum_n (SOD / L_n)^2."""
    import torch
    g = ctr.config_geometry("fdk_512")
    q = torch.ones((g.n_views, g.nv, g.nu), dtype=torch.float32, device=dev)
    staged, _ = ctr.stage_projections(q, g, ctr.F32)
    vol = ctr.backproject(staged, g, ctr.F32, scale=1.0 / g.n_views)
    torch.cuda.synchronize()
    d = g.voxel
    th = 2 * np.pi * np.arange(g.n_views) / g.n_views
    for j in (256, 100):
        sl = vol[:, j, :].cpu().numpy()           # (Z, X)
        k = np.arange(g.Z)[:, None, None]
        i = np.arange(g.X)[None, :, None]
        x, z = (i + 0.5 - g.X / 2) * d, (k + 0.5 - g.Z / 2) * d
        L = g.source_distance - x * np.cos(th) - z * np.sin(th)
        ref = ((g.source_distance / L) ** 2).mean(axis=2)
        r = np.hypot(x[..., 0], z[..., 0])
        inside = r < 120.0                         # stays on the panel for every view
        assert np.abs(sl[inside] - ref[inside]).max() / ref[inside].max() < 2e-5


def test_full_size_filter_rows_sample(ctr, orc, dev):
    """FDK weight + 2048-point ramp filter on the 512^3 scan: GPU staged
    output vs the oracle on a sample of views."""
    import torch
    g = ctr.config_geometry("fdk_512")
    p = _random_views(ctr, g, dev, 7).abs_()
    staged, _ = ctr.filter_projections(p, g, ctr.RAM_LAK, 2048, ctr.F32, fdk=True)
    pu, pv, _ = ctr.staged_layout(g.nu, g.nv, g.n_views)
    st = staged.view(torch.float32).view(g.n_views, pv, pu)
    ph = p.cpu().numpy()
    for n in (0, 359, 719):
        got = st[n, 1:g.nv + 1, 1:g.nu + 1].cpu().numpy()
        w = orc.fdk_weight(ph[n:n + 1], g.source_distance, g.detector_distance,
                           g.detector_width)[0]
        ref = orc.filter_rows(w.astype(np.float32), orc.RAM_LAK, 2048)
        assert rel_max(got, ref) < 1e-5


def test_backproject_kernel_variants_agree(ctr, dev):
    """The LDS-staged kernel and the direct-tap kernel agree over the whole
    256^3 volume, including the corner tiles whose projections leave the
    detector sideways in some views (zero texels there).  Both kernels take
    their coordinates from the same FP64 column set-up, interpolate in FP32
    the same way, and the LDS footprint holds exactly the taps the direct
    kernel reads; only the y-run length (32 vs 16 voxels) changes where the
    FP32 row step starts, so a sample's row coordinate differs by FP32
    rounding.  FP32 data then agree to FP32 rounding (about 2e-6 of the
    peak).  The 16-bit value-and-step texels are linear within a cell but
    not exactly continuous across cell boundaries: the lower tap of cell k
    is value + step of row k, which differs from row k + 1's own value by
    the step's rounding (at most half a 16-bit step of the row difference).
    A sample whose row coordinate lies within FP32 rounding of an integer
    can fall on either side, so the kernels agree to that rounding there:
    measured on white-noise views (row differences as large as the values)
    about 2e-4 max-rel for BF16 and a quarter of it for FP16, with the
    RMS difference at 3e-6 and below.  The LDS kernel of FP16 also rounds
    its row coordinates to 2^-16 (a row held biased by 128 gives its LDS
    address in one instruction): a uniform rounding of 4.4e-6 rows RMS,
    times white-noise row differences of about 1.4 times the values, is an
    RMS difference of about 6e-6 of the values, views being independent
    (measured 1.1e-5 with the direct kernel's own rounding)."""
    import torch
    g = ctr.config_geometry("fdk_256_s")
    q = _random_views(ctr, g, dev, 5)
    lds, gather = ctr.Workspace(), ctr.Workspace()
    lds.set_option(ctr.OPTION_BACKPROJECT_KERNEL, ctr.KERNEL_LDS)
    gather.set_option(ctr.OPTION_BACKPROJECT_KERNEL, ctr.KERNEL_GATHER)
    for dt in (ctr.F32, ctr.BF16, ctr.F16):
        staged, vs = ctr.stage_projections(q, g, dt)
        a = ctr.backproject(staged, g, dt, view_scale=vs, scale=0.01, ws=lds)
        b = ctr.backproject(staged, g, dt, view_scale=vs, scale=0.01, ws=gather)
        torch.cuda.synchronize()
        a, b = a.cpu().numpy(), b.cpu().numpy()
        print(f"[variants] dtype {dt}: max {rel_max(a, b):.3e} "
              f"rmse {rel_rmse(a, b):.3e}", flush=True)
        bound = {ctr.F32: (5e-6, 5e-6), ctr.F16: (1e-4, 3e-5), ctr.BF16: (5e-4, 1e-5)}[dt]
        assert rel_max(a, b) < bound[0], dt
        assert rel_rmse(a, b) < bound[1], dt


def test_lds_kernel_falls_back_on_wide_footprints(ctr, orc, dev):
    """High magnification: tile footprints wider than the LDS row take the
    direct-tap path inside the LDS kernel; parity is unchanged."""
    import torch
    g = ctr.Geometry(ctr.CONE, nu=512, nv=512, n_views=60, X=64, Y=64, Z=64,
                     detector_width=640.0, volume_width=32.0, source_distance=40.0,
                     detector_distance=400.0)
    rng = np.random.default_rng(17)
    qh = rng.standard_normal((g.n_views, g.nv, g.nu)).astype(np.float32)
    q = torch.from_numpy(qh).to(dev)
    staged, _ = ctr.stage_projections(q, g, ctr.F32)
    ws = ctr.Workspace()
    ws.set_option(ctr.OPTION_BACKPROJECT_KERNEL, ctr.KERNEL_LDS)
    vol = ctr.backproject(staged, g, ctr.F32, scale=0.01, ws=ws).cpu().numpy()
    box = ((20, 36), (0, 64), (20, 36))
    ref = orc.backproject_cone_region(qh, g.source_distance, g.detector_distance,
                                      g.detector_width, g.volume_width, 1.0, 0.0, 0.01,
                                      (g.X, g.Y, g.Z), box)
    assert rel_max(vol[20:36, :, 20:36], ref) < 1e-4


def test_full_size_forward_kernels_agree(ctr, dev):
    """The benchmark's projections (30 ellipsoids in 512^3, 720 views of
    768^2) from the slab-staged forward projector and from the ray kernel:
    identical bit for bit at the real shape."""
    import torch
    g = ctr.config_geometry("fdk_512")
    ell = ctr.random_ellipsoids(20240602, 30, (g.X, g.Y, g.Z), g.voxel, (5, 120))
    vol = torch.zeros((g.Z, g.Y, g.X), dtype=torch.float32, device=dev)
    ctr.render_ellipsoids(ell, vol, g.volume_width)
    slab_ws, ray_ws = ctr.Workspace(), ctr.Workspace()
    slab_ws.set_option(ctr.OPTION_FORWARD_KERNEL, ctr.KERNEL_AUTO)
    ray_ws.set_option(ctr.OPTION_FORWARD_KERNEL, ctr.KERNEL_RAY)
    slab = ctr.forward_project(vol, g, step=1.0, ws=slab_ws)
    ray = ctr.forward_project(vol, g, step=1.0, ws=ray_ws)
    torch.cuda.synchronize()
    assert float(ray.abs().max()) > 0.0
    assert torch.equal(slab, ray)


@pytest.mark.parametrize("Y", [64, 8])
def test_2048_panels_vs_oracle(ctr, orc, dev, Y):
    """2048 x 2048 panels, the reference's native scan frames
    (Mains/fbp_cone_fdk_main.cpp:23): a staged view is 2080 x 2051 texels
    (17 MB), past 2^22 texels.  A 2048 x Y x 2048 slab at 0.1 mm (1 detector
    pixel per voxel at magnification 2) from 16 views, Y = 64 through the
    LDS kernel and Y = 8 through the direct-tap kernel, FP32 and FP16
    storage, against the FP64 oracle on crops at the centre and near a
    corner of the field."""
    import torch
    g = ctr.Geometry(ctr.CONE, nu=2048, nv=2048, n_views=16, X=2048, Y=Y, Z=2048,
                     detector_width=409.6, volume_width=204.8, source_distance=500.0,
                     detector_distance=500.0, rotation_offset=0.1)
    gen = torch.Generator(device=dev)
    gen.manual_seed(2048 + Y)
    q = torch.randn((g.n_views, g.nv, g.nu), generator=gen, device=dev)
    qh = q.cpu().numpy()
    c = min(16, Y)
    boxes = (((1016, 1032), (Y // 2 - c // 2, Y // 2 + c // 2), (1016, 1032)),
             ((120, 136), (0, c), (1890, 1906)))
    refs = [orc.backproject_cone_region(qh, g.source_distance, g.detector_distance,
                                        g.detector_width, g.volume_width, 1.0, 0.1, 0.01,
                                        (g.X, g.Y, g.Z), box) for box in boxes]
    for dt in (ctr.F32, ctr.F16):
        staged, vs = ctr.stage_projections(q, g, dt)
        vol = ctr.backproject(staged, g, dt, view_scale=vs, scale=0.01)
        for box, ref in zip(boxes, refs):
            got = _crop(vol, box)
            assert np.abs(ref).max() > 0
            if dt == ctr.F32:
                assert rel_max(got, ref) < 1e-4, (Y, box)
            else:
                assert rel_rmse(got, ref) < 1e-3, (Y, box)
        del staged, vol
        torch.cuda.synchronize()
