"""The 16-bit staged forms (FP16 / BF16 value-and-step texels), through the
C ABI.

A staged 16-bit texel at (column c, row r) of view n holds the value
v = q[r][c] / view_scale[n] rounded to 16 bits in its lower half and, in its
upper half, the step d to the next staged row, q[r + 1][c] / view_scale[n] - v
(FP16: rounded to FP16; BF16: the upper half chosen so that the whole word,
read as FP32, is nearest to d).  The backprojector's vertical taps of a
column are v and v + d, interpolated and accumulated in FP32.

Checks:
* the staged texels themselves, decoded on the host, bit for bit against
  the same rounding done in numpy / torch (values, steps, borders);
* the backprojection of 16-bit staged data against the FP64 oracle of the
  DEQUANTIZED texels (oracle.backproject_*_cells: each interpolation cell
  sees its own top and bottom rows, v and v + d), which takes the storage
  rounding out, for the LDS kernel, the direct-tap kernel and parallel
  beams.  Within a cell the texels are exactly linear, so FP32 arithmetic
  gives FP32 rounding: rel RMSE <= 1e-5 for the direct-tap kernel (16-bit
  interpolation arithmetic would show 1e-4 and more on these white-noise
  views).  The LDS kernel of FP16 holds its row coordinates biased by
  128, rounded to 2^-16 of a row (a texture sampler's weights carry 2^-8):
  on white noise, whose row differences are as large as the values, that
  is a rel RMSE of about 1.1e-5, bounded at 2.5e-5.  Across a cell
  boundary the lower tap of cell k (v + d of row k) differs from row
  k + 1's own value by the step's rounding, so a sample whose FP32 row
  coordinate lies within rounding of an integer can land in the other cell
  than the FP64 oracle puts it in: the max-rel bound is the FP32 bar 1e-4
  for FP16 and 4e-4 for BF16, whose steps carry 8 bits;
* the staging ignores what the workspace's FP32 buffer or the caller's FP32
  staged buffer held before (borders included);
* packed launches equal single launches bit for bit;
* per-workspace options hold per thread: two threads with two workspaces and
  two kernel choices each reproduce their own result.
"""
import numpy as np
import pytest

from conftest import rel_max, rel_rmse

pytestmark = pytest.mark.gpu

# max-rel bounds against the dequantized oracle, and rel RMSE bounds per
# kernel (module docstring)
MAX_REL = {1: 1e-4, 2: 4e-4}
RMSE = {(1, "lds"): 2.5e-5, (1, "gather"): 1e-5, (2, "lds"): 1e-5, (2, "gather"): 1e-5}


def _cone(ctr, **kw):
    base = dict(kind=ctr.CONE, nu=64, nv=48, n_views=90, X=40, Y=32, Z=36,
                detector_width=128.0, volume_width=40.0, source_distance=150.0,
                detector_distance=100.0)
    base.update(kw)
    return ctr.Geometry(**base)


def texels(ctr, staged, g, dt):
    """The staged 16-bit texels as uint32 (n_views, pitch_v, pitch_u)."""
    import torch
    pu, pv, _ = ctr.staged_layout(g.nu, g.nv, g.n_views, dt)
    w = staged.view(torch.int32).view(g.n_views, pv, pu).cpu().numpy()
    return w.view(np.uint32), pu, pv


def halves_as_f32(w, dt):
    """(value, step) of value-and-step texels as exact float32 arrays."""
    lo = (w & 0xFFFF).astype(np.uint16)
    if dt == 1:  # CTR_F16
        hi = (w >> 16).astype(np.uint16)
        return lo.view(np.float16).astype(np.float32), hi.view(np.float16).astype(np.float32)
    val = (lo.astype(np.uint32) << 16).view(np.float32)
    return val, w.view(np.float32)


def dequantized_cells(ctr, staged, vs, g, dt):
    """Dequantized (top, bot) cells (n_views, nv + 1, nu) of a 16-bit staged
    buffer: cell c covers staged rows c and c + 1 (detector rows c - 1 and
    c); top = v * view_scale, bot = (v + d) * view_scale, in FP64 then FP32
    (the oracle's input type)."""
    w, pu, pv = texels(ctr, staged, g, dt)
    val, step = halves_as_f32(w, dt)
    s = vs.cpu().numpy().astype(np.float64)[:, None, None]
    v = val.astype(np.float64)[:, 0:g.nv + 1, 1:g.nu + 1]
    d = step.astype(np.float64)[:, 0:g.nv + 1, 1:g.nu + 1]
    return (v * s).astype(np.float32), ((v + d) * s).astype(np.float32)


def _f16(x):
    return np.asarray(x, dtype=np.float32).astype(np.float16)


def _bf16_bits(x):
    """Round-to-nearest-even float32 -> bf16 bits (uint32 in the low half)."""
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint32)


@pytest.mark.parametrize("dt", [1, 2])
def test_staged_texels_hold_value_and_step(ctr, dev, dt):
    """ctr_stage_projections writes value-and-step texels: the value is the
    16-bit rounding of q / view_scale, the step (FP16) the FP16 rounding of
    q[next row] / view_scale - value, or (BF16) a word whose FP32 reading is
    at least as near to that difference as its neighbours one bf16 step
    above and below; the border row holds (0, first data row); borders are
    zero otherwise; view_scale is max |q| per view."""
    import torch
    g = _cone(ctr, nu=70, nv=37, n_views=5)
    gen = torch.Generator(device=dev)
    gen.manual_seed(5 + dt)
    p = torch.randn((g.n_views, g.nv, g.nu), generator=gen, device=dev) * 3.0
    p[2] = 0.0  # an all-zero view: scale 1, all texels zero
    staged, vs = ctr.stage_projections(p, g, dt)
    torch.cuda.synchronize()
    ph = p.cpu().numpy()
    vsh = vs.cpu().numpy()
    ref_scale = np.abs(ph).reshape(g.n_views, -1).max(axis=1)
    ref_scale[ref_scale == 0] = 1.0
    assert np.array_equal(vsh, ref_scale.astype(np.float32))
    w, pu, pv = texels(ctr, staged, g, dt)
    # staged FP32 plane q (borders zero), scaled as the kernel does: q * (1 / s)
    q = np.zeros((g.n_views, pv + 1, pu), dtype=np.float32)
    q[:, 1:g.nv + 1, 1:g.nu + 1] = ph
    inv = (np.float32(1.0) / vsh.astype(np.float32))[:, None, None]
    qs = (q * inv).astype(np.float32)
    cur, nxt = qs[:, :pv], qs[:, 1:pv + 1]
    if dt == 1:
        v = _f16(cur)
        d = _f16(nxt - v.astype(np.float32))
        want = v.view(np.uint16).astype(np.uint32) | (d.view(np.uint16).astype(np.uint32) << 16)
        assert np.array_equal(w, want)
    else:
        lo = _bf16_bits(cur)
        assert np.array_equal(w & 0xFFFF, lo)
        t = (nxt - (lo << 16).view(np.float32)).astype(np.float64)
        got = np.abs(w.view(np.float32).astype(np.float64) - t)
        for delta in (-1, 1):
            alt = (((w >> 16) + np.uint32(delta & 0xFFFFFFFF)) & 0xFFFF) << 16 | (w & 0xFFFF)
            af = alt.view(np.float32).astype(np.float64)
            e = np.where(np.isfinite(af), np.abs(af - t), np.inf)
            assert (got <= e).all(), delta
    val, step = halves_as_f32(w, dt)
    assert (val[:, 0] == 0).all() and (val[:, g.nv + 1:] == 0).all()
    assert (step[:, g.nv + 1:] == 0).all()
    assert (val[:, :, 0] == 0).all() and (step[:, :, 0] == 0).all()
    assert (val[:, :, g.nu + 1:] == 0).all() and (step[:, :, g.nu + 1:] == 0).all()
    assert (w[2] == 0).all()
    # the border row's step is the first data row
    assert np.abs(step[:, 0, 1:g.nu + 1] - cur[:, 1, 1:g.nu + 1]).max() <= \
        2.0 ** -8 * np.abs(cur[:, 1]).max()


@pytest.mark.parametrize("X,Y,Z,n_views", [(40, 32, 36, 90), (37, 45, 29, 130),
                                           (64, 64, 64, 90), (33, 20, 31, 70)])
def test_backprojection_vs_dequantized_oracle(ctr, orc, dev, X, Y, Z, n_views):
    """FP32 interpolation of the 16-bit texels: the backprojection equals the
    FP64 oracle of the dequantized cells to FP32 rounding (module docstring),
    with the LDS kernel (Y >= 32 by default, forced for Y = 20) and the
    direct-tap kernel."""
    import torch
    g = _cone(ctr, X=X, Y=Y, Z=Z, n_views=n_views, rotation_offset=0.2)
    gen = torch.Generator(device=dev)
    gen.manual_seed(X * Z + n_views)
    p = torch.randn((g.n_views, g.nv, g.nu), generator=gen, device=dev)
    kernels = {}
    for name, k in (("lds", ctr.KERNEL_LDS), ("gather", ctr.KERNEL_GATHER)):
        kernels[name] = ctr.Workspace()
        kernels[name].set_option(ctr.OPTION_BACKPROJECT_KERNEL, k)
    for dt in (ctr.F16, ctr.BF16):
        staged, vs = ctr.stage_projections(p, g, dt)
        top, bot = dequantized_cells(ctr, staged, vs, g, dt)
        ref = orc.backproject_cone_cells(top, bot, g.source_distance, g.detector_distance,
                                         g.detector_width, g.volume_width, g.rotation_count,
                                         g.rotation_offset, 0.01, (g.X, g.Y, g.Z))
        for name, ws in kernels.items():
            got = ctr.backproject(staged, g, dt, view_scale=vs, scale=0.01,
                                  ws=ws).cpu().numpy()
            e, r = rel_max(got, ref), rel_rmse(got, ref)
            print(f"[storage16] {X}x{Y}x{Z}x{n_views} dt={dt} {name}: max-rel {e:.2e} "
                  f"rel-rmse {r:.2e}", flush=True)
            assert e < MAX_REL[dt] and r < RMSE[dt, name], (dt, name, e, r)


def test_tall_footprints_vs_dequantized_oracle(ctr, orc, dev):
    """Detector rows of 0.8 voxel at a magnification of 1.67 .. 1.92: a
    32-voxel run spans 67 .. 77 rows, so the LDS kernel's rectangles reach
    the top of its 72-row buffer (FP16 rows up to 71.95, biased to 199.95)
    and the taller ones take the direct-tap pass.  Both kernels against the
    FP64 oracle of the dequantized cells, as above."""
    import torch
    g = _cone(ctr, nu=112, nv=160, X=40, Y=64, Z=36, detector_width=89.6,
              rotation_offset=0.4)
    mag = (g.source_distance + g.detector_distance) / g.source_distance
    run_rows = 32 * (g.volume_width / g.X) * mag / (g.detector_width / g.nu)
    assert run_rows > 64
    gen = torch.Generator(device=dev)
    gen.manual_seed(123)
    p = torch.randn((g.n_views, g.nv, g.nu), generator=gen, device=dev)
    kernels = {}
    for name, k in (("lds", ctr.KERNEL_LDS), ("gather", ctr.KERNEL_GATHER)):
        kernels[name] = ctr.Workspace()
        kernels[name].set_option(ctr.OPTION_BACKPROJECT_KERNEL, k)
    for dt in (ctr.F16, ctr.BF16):
        staged, vs = ctr.stage_projections(p, g, dt)
        top, bot = dequantized_cells(ctr, staged, vs, g, dt)
        ref = orc.backproject_cone_cells(top, bot, g.source_distance, g.detector_distance,
                                         g.detector_width, g.volume_width, g.rotation_count,
                                         g.rotation_offset, 0.01, (g.X, g.Y, g.Z))
        for name, ws in kernels.items():
            got = ctr.backproject(staged, g, dt, view_scale=vs, scale=0.01,
                                  ws=ws).cpu().numpy()
            e, r = rel_max(got, ref), rel_rmse(got, ref)
            print(f"[storage16] tall footprints dt={dt} {name}: max-rel {e:.2e} "
                  f"rel-rmse {r:.2e}", flush=True)
            assert e < MAX_REL[dt] and r < RMSE[dt, name], (dt, name, e, r)


def test_parallel_backprojection_vs_dequantized_oracle(ctr, orc, dev):
    """Parallel beams (LDS kernel: 40 voxels along y; and the 2D direct-tap
    kernel on a one-row detector) against the oracle of the dequantized
    cells."""
    import torch
    cases = (ctr.Geometry(ctr.PARALLEL, nu=64, nv=48, n_views=100, X=40, Y=40, Z=36,
                          detector_width=64.0, volume_width=40.0, rotation_count=0.5,
                          rotation_offset=0.15),
             ctr.Geometry(ctr.PARALLEL, nu=91, nv=1, n_views=120, X=64, Y=1, Z=64,
                          detector_width=91.0, volume_width=64.0, rotation_count=0.5))
    for q, g in enumerate(cases):
        gen = torch.Generator(device=dev)
        gen.manual_seed(77 + q)
        p = torch.randn((g.n_views, g.nv, g.nu), generator=gen, device=dev)
        for dt in (ctr.F16, ctr.BF16):
            staged, vs = ctr.stage_projections(p, g, dt)
            top, bot = dequantized_cells(ctr, staged, vs, g, dt)
            ref = orc.backproject_parallel_cells(top, bot, g.detector_width, g.volume_width,
                                                 g.rotation_count, g.rotation_offset, 0.02,
                                                 (g.X, g.Y, g.Z))
            got = ctr.backproject(staged, g, dt, view_scale=vs, scale=0.02).cpu().numpy()
            e, r = rel_max(got, ref), rel_rmse(got, ref)
            print(f"[storage16] parallel {g.X}x{g.Y}x{g.Z} dt={dt}: max-rel {e:.2e} "
                  f"rel-rmse {r:.2e}", flush=True)
            assert e < MAX_REL[dt] and r < 1e-5, (q, dt, e, r)


def test_16bit_forms_near_fp32_on_smooth_data(ctr, orc, dev):
    """Against the FP64 oracle of the unrounded FP32 data, the 16-bit forms
    stay within their storage rounding: FP16 and BF16 rel RMSE of a
    filtered analytic-ball scan under 1e-3 (the 16-bit gate)."""
    import torch
    g = _cone(ctr, X=64, Y=64, Z=64, nu=96, nv=80, detector_width=180.0, volume_width=64.0,
              source_distance=200.0, detector_distance=150.0, n_views=120)
    ball = np.array([[0.0, 0.0, 0.0, 20.0, 20.0, 20.0, 0.0, 0.0, 0.0, 1.0]])
    proj = orc.analytic_cone(ball, g.source_distance, g.detector_distance, g.detector_width,
                             g.nu, g.nv, g.n_views, g.rotation_count,
                             g.rotation_offset).astype(np.float32)
    p = torch.from_numpy(proj).to(dev)
    nfft = 256
    scale = g.fbp_scale(nfft)
    staged32, _ = ctr.filter_projections(p, g, ctr.RAM_LAK, nfft, ctr.F32, fdk=True)
    pu, pv, _ = ctr.staged_layout(g.nu, g.nv, g.n_views)
    filt = staged32.view(torch.float32).view(g.n_views, pv, pu)[:, 1:g.nv + 1,
                                                                 1:g.nu + 1].cpu().numpy()
    ref = orc.backproject_cone(filt, g.source_distance, g.detector_distance, g.detector_width,
                               g.volume_width, g.rotation_count, g.rotation_offset, scale,
                               (g.X, g.Y, g.Z))
    for dt, name in ((ctr.F16, "fp16"), (ctr.BF16, "bf16")):
        staged, vs = ctr.filter_projections(p, g, ctr.RAM_LAK, nfft, dt, fdk=True)
        got = ctr.backproject(staged, g, dt, view_scale=vs, scale=scale).cpu().numpy()
        e = float(np.linalg.norm(got - ref) / np.linalg.norm(ref))
        print(f"[storage16] ball {name}: rel RMSE {e:.2e}", flush=True)
        assert e < 1e-3, (name, e)


@pytest.mark.parametrize("dt", [0, 1, 2])
def test_filter_staging_ignores_old_buffer_contents(ctr, dev, dt):
    """The filter writes the detector texels of its FP32 staging buffer only.
    A workspace whose staging buffer last held a larger scan (non-zero where
    this scan's border lies), or an FP32 staged buffer full of garbage, gives
    the same staged texels and scales as fresh buffers, borders included."""
    import torch
    big = _cone(ctr, nu=100, nv=70, n_views=40)
    g = _cone(ctr, nu=70, nv=37, n_views=30)
    gen = torch.Generator(device=dev)
    gen.manual_seed(11 + dt)
    pb = torch.rand((big.n_views, big.nv, big.nu), generator=gen, device=dev) + 1.0
    p = torch.randn((g.n_views, g.nv, g.nu), generator=gen, device=dev)
    used = ctr.Workspace()
    ctr.filter_projections(pb, big, ctr.RAM_LAK, 256, dt, ws=used)
    _, _, nbytes = ctr.staged_layout(g.nu, g.nv, g.n_views, dt)
    dirty = torch.full((nbytes // 4,), 7.0, device=dev).view(torch.uint8)
    got, vs = ctr.filter_projections(p, g, ctr.RAM_LAK, 256, dt, ws=used, staged=dirty)
    ref, vs_ref = ctr.filter_projections(p, g, ctr.RAM_LAK, 256, dt, ws=ctr.Workspace())
    torch.cuda.synchronize()
    assert torch.equal(got, ref)
    if dt:
        assert torch.equal(vs, vs_ref)
    w = got.view(torch.int32).view(g.n_views, *ctr.staged_layout(g.nu, g.nv, g.n_views, dt)[1::-1])
    w = w.cpu().numpy()
    if dt == 0:
        assert (w[:, 0] == 0).all() and (w[:, g.nv + 1:] == 0).all()
    else:  # the first row holds (0, first data row): its value half is zero
        assert (w[:, 0] & 0xFFFF == 0).all() and (w[:, g.nv + 1:] == 0).all()
    assert (w[:, :, 0] == 0).all() and (w[:, :, g.nu + 1:] == 0).all()


def test_packed_16bit_equals_single(ctr, dev):
    """Packed launches of 16-bit jobs are bit-identical to single launches."""
    import torch
    gs = [_cone(ctr), _cone(ctr, X=64, Y=64, Z=64, nu=96, nv=80, detector_width=180.0,
                            volume_width=64.0, source_distance=200.0,
                            detector_distance=150.0)]
    for dt in (ctr.F16, ctr.BF16):
        items, singles = [], []
        for q, g in enumerate(gs):
            gen = torch.Generator(device=dev)
            gen.manual_seed(40 + q)
            p = torch.randn((g.n_views, g.nv, g.nu), generator=gen, device=dev)
            staged, vs = ctr.stage_projections(p, g, dt)
            vol = torch.empty((g.Z, g.Y, g.X), device=dev)
            items.append((staged, g, vs, 0.02, vol, False))
            singles.append(ctr.backproject(staged, g, dt, view_scale=vs, scale=0.02))
        ctr.backproject_packed(items, dt)
        torch.cuda.synchronize()
        for it, ref in zip(items, singles):
            assert torch.equal(it[4], ref), dt


def test_two_threads_two_workspaces_concurrently(ctr, dev):
    """Two threads, each with its own stream and workspace, backproject the
    same FP16 data at the same time, one with the LDS kernel and one with
    the direct-tap kernel: each result equals its workspace's result made
    alone, bit for bit (options are per workspace, not per process)."""
    import threading
    import torch
    g = _cone(ctr, X=64, Y=64, Z=64, nu=96, nv=80, detector_width=180.0, volume_width=64.0,
              source_distance=200.0, detector_distance=150.0)
    gen = torch.Generator(device=dev)
    gen.manual_seed(91)
    p = torch.randn((g.n_views, g.nv, g.nu), generator=gen, device=dev)
    staged, vs = ctr.stage_projections(p, g, ctr.F16)
    wss = {k: ctr.Workspace() for k in (ctr.KERNEL_LDS, ctr.KERNEL_GATHER)}
    for k, w in wss.items():
        w.set_option(ctr.OPTION_BACKPROJECT_KERNEL, k)
        assert w.get_option(ctr.OPTION_BACKPROJECT_KERNEL) == k
    alone = {k: ctr.backproject(staged, g, ctr.F16, view_scale=vs, scale=0.02, ws=w)
             for k, w in wss.items()}
    torch.cuda.synchronize()
    results = {k: [] for k in wss}
    errors = []
    barrier = threading.Barrier(len(wss))

    def worker(k):
        try:
            s = torch.cuda.Stream()
            barrier.wait()
            with torch.cuda.stream(s):
                for _ in range(6):
                    results[k].append(ctr.backproject(staged, g, ctr.F16, view_scale=vs,
                                                      scale=0.02, ws=wss[k], stream=s))
            s.synchronize()
        except Exception as e:  # reported by the main thread
            errors.append(e)

    threads = [threading.Thread(target=worker, args=(k,)) for k in wss]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=60)
    assert not errors, errors
    for k in wss:
        assert len(results[k]) == 6
        for r in results[k]:
            assert torch.equal(r, alone[k]), k
