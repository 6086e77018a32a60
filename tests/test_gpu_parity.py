"""HIP path vs the FP64 CPU oracle, through the C ABI, on seeded inputs.

Tolerances: FP32 kernels <= 1e-4 max relative error (max |gpu - oracle| /
max |oracle|); FP16-stored projections <= 1e-3 relative RMSE; BF16-stored
projections <= 4e-3 relative RMSE (8-bit mantissa, reported alongside).
"""
import numpy as np
import pytest

from conftest import rel_max, rel_rmse

pytestmark = pytest.mark.gpu


def _t(dev, a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).to(dev)


@pytest.mark.parametrize("n", [1, 2, 8, 64, 512, 2048, 8192])
def test_fft_rows_vs_oracle(ctr, orc, dev, n):
    import torch
    rng = np.random.default_rng(n)
    rows = 5
    z = (rng.standard_normal((rows, n)) + 1j * rng.standard_normal((rows, n))).astype(np.complex64)
    t = torch.from_numpy(z).to(dev)
    ctr.fft_radix2(t)
    got = t.cpu().numpy()
    ref = np.stack([orc.fft_radix2(r) for r in z.astype(np.complex128)])
    assert rel_max(got, ref) < 2e-6 * max(1, np.log2(n))
    ctr.fft_radix2(t, inverse=True)
    assert rel_max(t.cpu().numpy(), z) < 1e-5


def test_fft_rejects_non_power_of_two(ctr, dev):
    import torch
    t = torch.zeros((2, 12), dtype=torch.complex64, device=dev)
    with pytest.raises(ctr.CtrError) as e:
        ctr.fft_radix2(t)
    assert e.value.code == -2


@pytest.mark.parametrize("width,npad,ftype", [
    (367, 0, 0), (367, 1024, 2), (768, 2048, 0), (768, 2048, 1), (768, 2048, 2),
    (100, 128, 0), (1, 1, 0), (5, 8, 2), (1536, 4096, 0),
    (512, 512, 0), (65, 1024, 1), (1000, 2048, 2), (2048, 2048, 0)])
def test_filter_rows_vs_oracle(ctr, orc, dev, width, npad, ftype):
    rng = np.random.default_rng(width + npad)
    rows = 7
    x = rng.standard_normal((rows, width)).astype(np.float32)
    got = ctr.filter_rows(_t(dev, x), ftype, npad, out=_t(dev, np.zeros_like(x))).cpu().numpy()
    ref = orc.filter_rows(x, ftype, npad or None)
    assert rel_max(got, ref) < 1e-5


def test_filter_rows_in_place_odd_rows(ctr, orc, dev):
    rng = np.random.default_rng(3)
    x = rng.standard_normal((9, 200)).astype(np.float32)
    t = _t(dev, x)
    ctr.filter_rows(t, 2, 512)
    assert rel_max(t.cpu().numpy(), orc.filter_rows(x, 2, 512)) < 1e-5


def _cone(ctr, **kw):
    base = dict(kind=ctr.CONE, nu=64, nv=48, n_views=90, X=40, Y=32, Z=36,
                detector_width=128.0, volume_width=40.0, source_distance=150.0,
                detector_distance=100.0)
    base.update(kw)
    return ctr.Geometry(**base)


@pytest.mark.parametrize("layout,nfft", [(0, 256), (1, 256), (0, 1024), (1, 1024)])
def test_filter_projections_fdk_vs_oracle(ctr, orc, dev, layout, nfft):
    import torch
    g = _cone(ctr)
    rng = np.random.default_rng(11)
    proj = rng.standard_normal((g.n_views, g.nv, g.nu)).astype(np.float32)
    src = proj if layout == 0 else np.ascontiguousarray(proj.transpose(1, 0, 2))
    staged, _ = ctr.filter_projections(_t(dev, src), g, 0, nfft, ctr.F32, fdk=True,
                                       layout=layout)
    pu, pv, _ = ctr.staged_layout(g.nu, g.nv, g.n_views)
    st = staged.view(torch.float32).cpu().numpy().reshape(g.n_views, pv, pu)
    assert np.all(st[:, 0, :] == 0) and np.all(st[:, g.nv + 1:, :] == 0)
    assert np.all(st[:, :, 0] == 0) and np.all(st[:, :, g.nu + 1:] == 0)
    got = st[:, 1:g.nv + 1, 1:g.nu + 1]
    w = orc.fdk_weight(proj, g.source_distance, g.detector_distance, g.detector_width)
    ref = orc.filter_rows(w.astype(np.float32), 0, nfft)
    assert rel_max(got, ref) < 1e-5


@pytest.mark.parametrize("dt", [1, 2])
def test_filter_projections_16bit_equals_staged_fp32(ctr, dev, dt):
    """16-bit staging by the filter (per-row maxima fused into the filter,
    then the per-view scale and the value-and-step texels) equals staging
    the filter's FP32 output afterwards (ctr_stage_projections), bit for
    bit; and both input layouts give the same staged buffers bit for bit
    (the filter pairs rows in projection order in either layout), with an
    all-zero view staying exactly zero (scale 1)."""
    import torch
    g = _cone(ctr, n_views=37)
    rng = np.random.default_rng(12 + dt)
    proj = rng.standard_normal((g.n_views, g.nv, g.nu)).astype(np.float32)
    proj[5] = 0.0  # an all-zero view
    pu, pv, _ = ctr.staged_layout(g.nu, g.nv, g.n_views)
    out = []
    for layout in (0, 1):
        src = _t(dev, proj if layout == 0 else np.ascontiguousarray(proj.transpose(1, 0, 2)))
        s32, _ = ctr.filter_projections(src, g, 0, 256, ctr.F32, fdk=True, layout=layout)
        q = s32.view(torch.float32).view(g.n_views, pv, pu)[:, 1:g.nv + 1, 1:g.nu + 1]
        want, want_vs = ctr.stage_projections(q.contiguous(), g, dt)
        got, vs = ctr.filter_projections(src, g, 0, 256, dt, fdk=True, layout=layout)
        torch.cuda.synchronize()
        assert torch.equal(vs, want_vs), layout
        assert torch.equal(got, want), layout
        assert float(vs[5]) == 1.0 and not q[5].any(), layout
        out.append((s32, got, vs))
    assert torch.equal(out[0][0], out[1][0])
    assert torch.equal(out[0][1], out[1][1]) and torch.equal(out[0][2], out[1][2])


def _fdk_ref(orc, g, q, scale):
    return orc.backproject_cone(q, g.source_distance, g.detector_distance,
                                g.detector_width, g.volume_width, g.rotation_count,
                                g.rotation_offset, scale, (g.X, g.Y, g.Z))


@pytest.mark.parametrize("shape", [(40, 32, 36), (17, 9, 23), (64, 64, 64)])
def test_backproject_cone_f32_vs_oracle(ctr, orc, dev, shape):
    X, Y, Z = shape
    g = _cone(ctr, X=X, Y=Y, Z=Z, rotation_offset=0.3)
    rng = np.random.default_rng(X)
    q = rng.standard_normal((g.n_views, g.nv, g.nu)).astype(np.float32)
    staged, _ = ctr.stage_projections(_t(dev, q), g, ctr.F32)
    got = ctr.backproject(staged, g, ctr.F32, scale=0.01).cpu().numpy()
    ref = _fdk_ref(orc, g, q, 0.01)
    assert rel_max(got, ref) < 1e-4


def test_backproject_accumulate_and_clip(ctr, orc, dev):
    import torch
    g = _cone(ctr)
    rng = np.random.default_rng(5)
    q = rng.standard_normal((g.n_views, g.nv, g.nu)).astype(np.float32)
    staged, _ = ctr.stage_projections(_t(dev, q), g, ctr.F32)
    base = torch.full((g.Z, g.Y, g.X), 0.25, device=dev)
    got = ctr.backproject(staged, g, ctr.F32, scale=0.01, volume=base, accumulate=True,
                          clip=True).cpu().numpy()
    ref = np.maximum(_fdk_ref(orc, g, q, 0.01) + 0.25, 0.0)
    assert rel_max(got, ref) < 1e-4
    assert got.min() >= 0.0


@pytest.mark.parametrize("dtype,tol", [(1, 1e-3), (2, 4e-3)])
def test_backproject_16bit_storage(ctr, orc, dev, dtype, tol):
    g = _cone(ctr)
    rng = np.random.default_rng(21)
    p = rng.standard_normal((g.n_views, g.nv, g.nu)).astype(np.float32)
    nfft = 256
    staged, vs = ctr.filter_projections(_t(dev, p), g, 0, nfft, dtype, fdk=True)
    assert vs is not None
    got = ctr.backproject(staged, g, dtype, view_scale=vs, scale=g.fbp_scale(nfft)).cpu().numpy()
    w = orc.fdk_weight(p, g.source_distance, g.detector_distance, g.detector_width)
    q = orc.filter_rows(w.astype(np.float32), 0, nfft).astype(np.float32)
    ref = _fdk_ref(orc, g, q, g.fbp_scale(nfft))
    assert rel_rmse(got, ref) < tol


def test_backproject_parallel_vs_oracle(ctr, orc, dev):
    g = ctr.Geometry(ctr.PARALLEL, nu=50, nv=20, n_views=64, X=32, Y=16, Z=30,
                     detector_width=60.0, volume_width=40.0, rotation_count=0.5,
                     rotation_offset=0.1)
    rng = np.random.default_rng(9)
    q = rng.standard_normal((g.n_views, g.nv, g.nu)).astype(np.float32)
    staged, _ = ctr.stage_projections(_t(dev, q), g, ctr.F32)
    got = ctr.backproject(staged, g, ctr.F32, scale=0.02).cpu().numpy()
    ref = orc.backproject_parallel(q, g.detector_width, g.volume_width, g.rotation_count,
                                   g.rotation_offset, 0.02, (g.X, g.Y, g.Z))
    assert rel_max(got, ref) < 1e-4


def test_sinogram_layout_matches_projection_layout(ctr, dev):
    g = _cone(ctr)
    rng = np.random.default_rng(2)
    q = rng.standard_normal((g.n_views, g.nv, g.nu)).astype(np.float32)
    a, _ = ctr.stage_projections(_t(dev, q), g, ctr.F32, layout=0)
    b, _ = ctr.stage_projections(_t(dev, q.transpose(1, 0, 2)), g, ctr.F32, layout=1)
    va = ctr.backproject(a, g, ctr.F32).cpu().numpy()
    vb = ctr.backproject(b, g, ctr.F32).cpu().numpy()
    assert np.array_equal(va, vb)


@pytest.mark.parametrize("kind", ["cone", "parallel"])
def test_forward_project_vs_oracle(ctr, orc, dev, kind):
    import torch
    rng = np.random.default_rng(4)
    X, Y, Z = 24, 20, 28
    vol = rng.random((Z, Y, X)).astype(np.float32)
    if kind == "cone":
        g = ctr.Geometry(ctr.CONE, nu=40, nv=36, n_views=24, X=X, Y=Y, Z=Z,
                         detector_width=80.0, volume_width=24.0, source_distance=60.0,
                         detector_distance=40.0, rotation_offset=0.2)
        ref = orc.forward_cone(vol, g.volume_width, g.source_distance, g.detector_distance,
                               g.detector_width, g.nu, g.nv, g.n_views, 1.0, 0.2, 0.5)
    else:
        g = ctr.Geometry(ctr.PARALLEL, nu=40, nv=24, n_views=24, X=X, Y=Y, Z=Z,
                         detector_width=40.0, volume_width=24.0, rotation_count=0.5)
        ref = orc.forward_parallel(vol, g.volume_width, g.detector_width, g.nu, g.nv,
                                   g.n_views, 0.5, 0.0, 0.5)
    got = ctr.forward_project(_t(dev, vol), g, step=0.5).cpu().numpy()
    assert rel_max(got, ref) < 1e-4
    sino = ctr.forward_project(_t(dev, vol), g, layout=ctr.SINOGRAMS, step=0.5)
    assert torch.equal(sino.permute(1, 0, 2).cpu(), torch.from_numpy(got))


@pytest.mark.parametrize("kind", ["cone", "parallel"])
def test_forward_project_interior_vs_oracle(ctr, orc, dev, kind):
    """A volume large enough that most samples of a ray lie inside it and
    take the unchecked interior loop (the samples near the faces still take
    the checked one): against the oracle's ray-marcher, and the slab-staged
    kernel (the default) and the ray kernel, each with buffer-resource
    addressing (below 4 GiB) and 64-bit addressing, bit-identical."""
    import torch
    rng = np.random.default_rng(41)
    X, Y, Z = 64, 48, 56
    vol = rng.random((Z, Y, X)).astype(np.float32)
    if kind == "cone":
        g = ctr.Geometry(ctr.CONE, nu=80, nv=64, n_views=30, X=X, Y=Y, Z=Z,
                         detector_width=160.0, volume_width=64.0, source_distance=150.0,
                         detector_distance=100.0, rotation_offset=0.1)
        ref = orc.forward_cone(vol, g.volume_width, g.source_distance, g.detector_distance,
                               g.detector_width, g.nu, g.nv, g.n_views, 1.0, 0.1, 0.5)
    else:
        g = ctr.Geometry(ctr.PARALLEL, nu=96, nv=48, n_views=30, X=X, Y=Y, Z=Z,
                         detector_width=96.0, volume_width=64.0, rotation_count=0.5)
        ref = orc.forward_parallel(vol, g.volume_width, g.detector_width, g.nu, g.nv,
                                   g.n_views, 0.5, 0.0, 0.5)
    got = ctr.forward_project(_t(dev, vol), g, step=0.5, ws=_forward_ws(ctr, False, False))
    assert rel_max(got.cpu().numpy(), ref) < 1e-4
    ray = ctr.forward_project(_t(dev, vol), g, step=0.5, ws=_forward_ws(ctr, True, False))
    assert torch.equal(ray, got)
    wide = ctr.forward_project(_t(dev, vol), g, step=0.5, ws=_forward_ws(ctr, True, True))
    assert torch.equal(wide, got)
    wide_slab = ctr.forward_project(_t(dev, vol), g, step=0.5,
                                    ws=_forward_ws(ctr, False, True))
    assert torch.equal(wide_slab, got)


def _forward_ws(ctr, ray, pointer):
    """Workspace choosing the forward kernel (ray or slab) and addressing
    (64-bit pointers or the buffer resource)."""
    ws = ctr.Workspace()
    ws.set_option(ctr.OPTION_FORWARD_KERNEL, ctr.KERNEL_RAY if ray else ctr.KERNEL_AUTO)
    ws.set_option(ctr.OPTION_FORWARD_ADDRESSING,
                  ctr.KERNEL_POINTER if pointer else ctr.KERNEL_AUTO)
    return ws


@pytest.mark.parametrize("step", [0.5, 1.0, 2.0])
def test_forward_slab_kernel_matches_ray_kernel(ctr, dev, step):
    """The slab-staged forward projector against the ray kernel, bit for bit:
    a full turn of views (every main axis and direction), a fan wide enough
    that some rays run against the view's central direction (those take the
    in-order fallback), a fan of +-85 degrees, odd sizes and several
    steps."""
    import torch
    rng = np.random.default_rng(int(step * 10))
    X, Y, Z = 45, 37, 51
    vol = _t(dev, rng.random((Z, Y, X)).astype(np.float32))
    geoms = [ctr.Geometry(ctr.CONE, nu=70, nv=50, n_views=37, X=X, Y=Y, Z=Z,
                          detector_width=140.0, volume_width=45.0, source_distance=120.0,
                          detector_distance=90.0, rotation_offset=0.05),
             ctr.Geometry(ctr.CONE, nu=90, nv=30, n_views=24, X=X, Y=Y, Z=Z,
                          detector_width=600.0, volume_width=45.0, source_distance=60.0,
                          detector_distance=40.0, rotation_offset=0.3),
             ctr.Geometry(ctr.PARALLEL, nu=64, nv=40, n_views=29, X=X, Y=Y, Z=Z,
                          detector_width=64.0, volume_width=45.0, rotation_count=0.5,
                          rotation_offset=0.2),
             # a fan of +-85 degrees (and +-62 degrees vertically): rays far
             # from the view's main axis, whose y / cross-axis cells test the
             # bounds of the staged box
             ctr.Geometry(ctr.CONE, nu=160, nv=48, n_views=23, X=X, Y=Y, Z=Z,
                          detector_width=2400.0, volume_width=45.0, source_distance=60.0,
                          detector_distance=40.0, rotation_offset=0.7)]
    for g in geoms:
        slab = ctr.forward_project(vol, g, step=step, ws=_forward_ws(ctr, False, False))
        slab_wide = ctr.forward_project(vol, g, step=step, ws=_forward_ws(ctr, False, True))
        ray = ctr.forward_project(vol, g, step=step, ws=_forward_ws(ctr, True, False))
        assert float(ray.abs().max()) > 0.0
        assert torch.equal(slab, ray), (g.kind, g.nu, step)
        assert torch.equal(slab_wide, ray), (g.kind, g.nu, step)


@pytest.mark.parametrize("n_views", [6, 24, 240])
def test_forward_slab_views_per_block_and_short_slabs(ctr, dev, n_views):
    """The slab kernel stages one box for 1 or 2 consecutive views, chosen by
    how far the volume's corner voxels move between views (37 voxels for 6
    views of this volume: one view a block; 9.3 and 0.93 for 24 and 240
    views: two), and shortens a slab whose
    box would not fit LDS (detector pixels about three voxels wide at the
    rotation axis here, so a 16 x 16 tile spans some 50 voxels).  Against the
    ray kernel bit for bit, with the box filled through the buffer resource
    and through 64-bit addresses."""
    import torch
    rng = np.random.default_rng(n_views)
    X, Y, Z = 48, 40, 52
    vol = _t(dev, rng.random((Z, Y, X)).astype(np.float32))
    g = ctr.Geometry(ctr.CONE, nu=56, nv=48, n_views=n_views, X=X, Y=Y, Z=Z,
                     detector_width=336.0, volume_width=48.0, source_distance=100.0,
                     detector_distance=100.0, rotation_offset=0.15)
    ray = ctr.forward_project(vol, g, step=1.0, ws=_forward_ws(ctr, True, False))
    slab = ctr.forward_project(vol, g, step=1.0, ws=_forward_ws(ctr, False, False))
    slab_wide = ctr.forward_project(vol, g, step=1.0, ws=_forward_ws(ctr, False, True))
    assert float(ray.abs().max()) > 0.0
    assert torch.equal(slab, ray), n_views
    assert torch.equal(slab_wide, ray), n_views


def test_fdk_end_to_end_ball(ctr, orc, dev):
    """Analytic projections of a uniform ball -> FDK recovers its density."""
    g = ctr.Geometry(ctr.CONE, nu=128, nv=128, n_views=180, X=64, Y=64, Z=64,
                     detector_width=256.0, volume_width=64.0, source_distance=200.0,
                     detector_distance=200.0)
    ball = np.array([[0, 0, 0, 20, 20, 20, 0, 0, 0, 0.02]], dtype=np.float64)
    p = orc.analytic_cone(ball, g.source_distance, g.detector_distance, g.detector_width,
                          g.nu, g.nv, g.n_views).astype(np.float32)
    nfft = 256
    staged, _ = ctr.filter_projections(_t(dev, p), g, 0, nfft, ctr.F32, fdk=True)
    vol = ctr.backproject(staged, g, ctr.F32, scale=g.fbp_scale(nfft)).cpu().numpy()
    centre = vol[28:36, 28:36, 28:36].mean()
    assert abs(centre - 0.02) < 0.02 * 0.03
    assert abs(vol[2:6, 30:34, 2:6].mean()) < 0.02 * 0.05


def test_error_paths(ctr, dev):
    import torch
    g = _cone(ctr)
    staged = torch.zeros(16, dtype=torch.uint8, device=dev)
    with pytest.raises(ctr.CtrError):
        ctr.backproject(staged, ctr.Geometry(**{**g.__dict__, "source_distance": 1.0}))
    with pytest.raises(ctr.CtrError):
        ctr.filter_rows(torch.zeros((2, 100), device=dev), 0, 96)
    with pytest.raises(ctr.CtrError):
        ctr.filter_rows(torch.zeros((2, 100), device=dev), 7, 128)


def test_lds_tile_order_covers_every_tile(ctr, dev):
    """The LDS kernel deals its tiles to the XCDs in chunks of 128 and walks
    bands of 4 tile rows inside each y-run layer.  On a grid of 11 x 13
    column tiles x 11 y-runs (1573 tiles: one whole round of 8 chunks and a
    tail that keeps its own order, a last band one tile row high, a partial
    last y-run) every voxel is written once: a NaN-filled volume comes back
    without NaN, equal to the direct-tap kernel to FP32 rounding."""
    import torch
    g = _cone(ctr, nu=320, nv=600, n_views=24, X=176, Y=330, Z=200,
              detector_width=320.0, volume_width=176.0, source_distance=400.0,
              detector_distance=300.0, rotation_offset=0.1)
    assert ((g.X + 15) // 16) * ((g.Z + 15) // 16) * ((g.Y + 31) // 32) == 1573
    gen = torch.Generator(device=dev)
    gen.manual_seed(21)
    q = torch.randn((g.n_views, g.nv, g.nu), generator=gen, device=dev)
    staged, _ = ctr.stage_projections(q, g, ctr.F32)
    lds, gather = ctr.Workspace(), ctr.Workspace()
    lds.set_option(ctr.OPTION_BACKPROJECT_KERNEL, ctr.KERNEL_LDS)
    gather.set_option(ctr.OPTION_BACKPROJECT_KERNEL, ctr.KERNEL_GATHER)
    vol = torch.full((g.Z, g.Y, g.X), float("nan"), device=dev)
    a = ctr.backproject(staged, g, ctr.F32, scale=0.01, volume=vol, ws=lds)
    b = ctr.backproject(staged, g, ctr.F32, scale=0.01, ws=gather)
    torch.cuda.synchronize()
    a, b = a.cpu().numpy(), b.cpu().numpy()
    assert np.isfinite(a).all()
    assert np.abs(b).max() > 0
    assert rel_max(a, b) < 5e-6
    # a packed launch over the same job and a small one (1573 + 18 tiles)
    # gives each job exactly its single launch
    g2 = _cone(ctr, X=40, Y=64, Z=36, rotation_offset=0.3)
    q2 = torch.randn((g2.n_views, g2.nv, g2.nu), generator=gen, device=dev)
    staged2, _ = ctr.stage_projections(q2, g2, ctr.F32)
    v1 = torch.full((g.Z, g.Y, g.X), float("nan"), device=dev)
    v2 = torch.full((g2.Z, g2.Y, g2.X), float("nan"), device=dev)
    ctr.backproject_packed([(staged, g, None, 0.01, v1, False),
                            (staged2, g2, None, 0.02, v2, True)], ctr.F32)
    s2 = ctr.backproject(staged2, g2, ctr.F32, scale=0.02, clip=True)
    torch.cuda.synchronize()
    assert np.array_equal(v1.cpu().numpy(), a)
    assert torch.equal(v2, s2)


@pytest.mark.parametrize("X,Y,Z,n_views", [(37, 45, 29, 130), (33, 64, 18, 1)])
def test_lds_kernel_edge_shapes(ctr, orc, dev, X, Y, Z, n_views):
    """The LDS-staged kernel on ragged tiles (X, Z not multiples of 16), a
    partial last y-run (Y not a multiple of 32), a view count that crosses
    the 128-view refill of the rectangle table, and a single view; every
    storage form against the FP64 oracle of the FP32 data."""
    g = _cone(ctr, X=X, Y=Y, Z=Z, n_views=n_views, rotation_offset=0.2)
    rng = np.random.default_rng(X * Y + n_views)
    q = rng.standard_normal((g.n_views, g.nv, g.nu)).astype(np.float32)
    ref = _fdk_ref(orc, g, q, 0.01)
    ws = ctr.Workspace()
    ws.set_option(ctr.OPTION_BACKPROJECT_KERNEL, ctr.KERNEL_LDS)
    for dt, tol in ((ctr.F32, 1e-4), (ctr.F16, 1e-3), (ctr.BF16, 4e-3)):
        staged, vs = ctr.stage_projections(_t(dev, q), g, dt)
        got = ctr.backproject(staged, g, dt, view_scale=vs, scale=0.01, ws=ws).cpu().numpy()
        if dt == ctr.F32:
            assert rel_max(got, ref) < tol, dt
        else:
            assert rel_rmse(got, ref) < tol, dt
