"""The two interpolation modes of the 16-bit staged forms, through the C ABI.

Under CTR_INTERP_FP32 (the default) the FP16 / BF16 pairs are widened and
interpolated in FP32, so the only difference from the FP32 path is the
storage rounding.
The reference here is the FP64 oracle backprojection of the DEQUANTIZED
staged data (each stored 16-bit texel times its view scale, exact in FP64),
which takes the storage rounding out: the FP32-interpolation result is held
to the FP32 bar, max-rel <= 1e-4.  The opt-in CTR_INTERP_FAST arithmetic is
checked against the same reference at the 16-bit forms' bars (FP16 1e-3,
BF16 4e-3 rel RMSE), which isolates the interpolation arithmetic's own
error, and must differ from the FP32-interpolation result (the switch takes
effect).  The packed launch follows the mode too and stays bit-identical to
single launches.
"""
import numpy as np
import pytest

from conftest import rel_max, rel_rmse

pytestmark = pytest.mark.gpu


def _cone(ctr, **kw):
    base = dict(kind=ctr.CONE, nu=64, nv=48, n_views=90, X=40, Y=32, Z=36,
                detector_width=128.0, volume_width=40.0, source_distance=150.0,
                detector_distance=100.0)
    base.update(kw)
    return ctr.Geometry(**base)


@pytest.fixture
def interp(ctr):
    """A workspace whose interpolation arithmetic the test sets."""
    return ctr.Workspace()


def dequantize(ctr, staged, vs, g, dt):
    """FP64 (n_views, nv, nu) values of a 16-bit staged buffer: the low half
    of the pair at staged (n, v + 1, u + 1) is q[n, v, u] / view_scale[n]."""
    import torch
    pu, pv, _ = ctr.staged_layout(g.nu, g.nv, g.n_views, dt)
    h = staged.view(torch.int16).view(g.n_views, pv, pu, 2)[..., 0].contiguous()
    h = h.view(torch.float16 if dt == ctr.F16 else torch.bfloat16).float().cpu().numpy()
    s = vs.cpu().numpy().astype(np.float64)
    return h[:, 1:g.nv + 1, 1:g.nu + 1].astype(np.float64) * s[:, None, None]


def _ref(orc, g, q, scale):
    return orc.backproject_cone(q, g.source_distance, g.detector_distance,
                                g.detector_width, g.volume_width, g.rotation_count,
                                g.rotation_offset, scale, (g.X, g.Y, g.Z))


@pytest.mark.parametrize("X,Y,Z,n_views", [(40, 32, 36, 90), (37, 45, 29, 130),
                                           (64, 64, 64, 90)])
def test_fp32_interpolation_vs_dequantized_oracle(ctr, orc, dev, interp, X, Y, Z, n_views):
    import torch
    g = _cone(ctr, X=X, Y=Y, Z=Z, n_views=n_views, rotation_offset=0.2)
    gen = torch.Generator(device=dev)
    gen.manual_seed(X * Z + n_views)
    p = torch.randn((g.n_views, g.nv, g.nu), generator=gen, device=dev)
    for dt, tol in ((ctr.F16, 1e-3), (ctr.BF16, 4e-3)):
        staged, vs = ctr.stage_projections(p, g, dt)
        ref = _ref(orc, g, dequantize(ctr, staged, vs, g, dt), 0.01)
        interp.set_interpolation(ctr.INTERP_FP32)
        strict = ctr.backproject(staged, g, dt, view_scale=vs, scale=0.01,
                                 ws=interp).cpu().numpy()
        interp.set_interpolation(ctr.INTERP_FAST)
        fast = ctr.backproject(staged, g, dt, view_scale=vs, scale=0.01,
                               ws=interp).cpu().numpy()
        e_strict, e_fast = rel_max(strict, ref), rel_rmse(fast, ref)
        print(f"[interp] {X}x{Y}x{Z}x{n_views} dt={dt}: fp32 max-rel {e_strict:.2e}, "
              f"fast rel-rmse {e_fast:.2e}", flush=True)
        assert e_strict < 1e-4, (dt, e_strict)
        assert e_fast < tol, (dt, e_fast)
        assert not np.array_equal(strict, fast), dt


def test_fp32_interpolation_parallel_vs_dequantized_oracle(ctr, orc, dev, interp):
    """Parallel-beam volume (the LDS kernel: 40 voxels along y) under both
    modes, against the FP64 oracle of the dequantized staged data."""
    import torch
    g = ctr.Geometry(ctr.PARALLEL, nu=64, nv=48, n_views=100, X=40, Y=40, Z=36,
                     detector_width=64.0, volume_width=40.0, rotation_count=0.5,
                     rotation_offset=0.15)
    gen = torch.Generator(device=dev)
    gen.manual_seed(77)
    p = torch.randn((g.n_views, g.nv, g.nu), generator=gen, device=dev)
    for dt, tol in ((ctr.F16, 1e-3), (ctr.BF16, 4e-3)):
        staged, vs = ctr.stage_projections(p, g, dt)
        q = dequantize(ctr, staged, vs, g, dt)
        ref = orc.backproject_parallel(q, g.detector_width, g.volume_width, g.rotation_count,
                                       g.rotation_offset, 0.02, (g.X, g.Y, g.Z))
        interp.set_interpolation(ctr.INTERP_FP32)
        strict = ctr.backproject(staged, g, dt, view_scale=vs, scale=0.02,
                                 ws=interp).cpu().numpy()
        interp.set_interpolation(ctr.INTERP_FAST)
        fast = ctr.backproject(staged, g, dt, view_scale=vs, scale=0.02,
                               ws=interp).cpu().numpy()
        e_strict, e_fast = rel_max(strict, ref), rel_rmse(fast, ref)
        print(f"[interp] parallel dt={dt}: fp32 max-rel {e_strict:.2e}, "
              f"fast rel-rmse {e_fast:.2e}", flush=True)
        assert e_strict < 1e-4, (dt, e_strict)
        assert e_fast < tol, (dt, e_fast)


def test_packed_follows_interpolation_mode(ctr, dev, interp):
    """Packed launches read the workspace's mode as single launches do, and
    stay bit-identical to them under both modes."""
    import torch
    gs = [_cone(ctr), _cone(ctr, X=64, Y=64, Z=64, nu=96, nv=80, detector_width=180.0,
                            volume_width=64.0, source_distance=200.0,
                            detector_distance=150.0)]
    for m in (ctr.INTERP_FP32, ctr.INTERP_FAST):
        interp.set_interpolation(m)
        for dt in (ctr.F16, ctr.BF16):
            items, singles = [], []
            for q, g in enumerate(gs):
                gen = torch.Generator(device=dev)
                gen.manual_seed(40 + q)
                p = torch.randn((g.n_views, g.nv, g.nu), generator=gen, device=dev)
                staged, vs = ctr.stage_projections(p, g, dt)
                vol = torch.empty((g.Z, g.Y, g.X), device=dev)
                items.append((staged, g, vs, 0.02, vol, False))
                singles.append(ctr.backproject(staged, g, dt, view_scale=vs, scale=0.02,
                                               ws=interp))
            ctr.backproject_packed(items, dt, ws=interp)
            torch.cuda.synchronize()
            for it, ref in zip(items, singles):
                assert torch.equal(it[4], ref), (m, dt)


def test_two_threads_two_modes_concurrently(ctr, dev):
    """Two threads, each with its own stream and workspace, backproject the
    same 16-bit data at the same time, one under FP32 interpolation and one
    under the 16-bit arithmetic: each result equals its mode's result made
    alone, bit for bit (the mode is per workspace, not per process)."""
    import threading
    import torch
    g = _cone(ctr, X=64, Y=64, Z=64, nu=96, nv=80, detector_width=180.0, volume_width=64.0,
              source_distance=200.0, detector_distance=150.0)
    gen = torch.Generator(device=dev)
    gen.manual_seed(91)
    p = torch.randn((g.n_views, g.nv, g.nu), generator=gen, device=dev)
    staged, vs = ctr.stage_projections(p, g, ctr.F16)
    wss = {m: ctr.Workspace() for m in (ctr.INTERP_FP32, ctr.INTERP_FAST)}
    for m, w in wss.items():
        w.set_interpolation(m)
    alone = {m: ctr.backproject(staged, g, ctr.F16, view_scale=vs, scale=0.02, ws=w)
             for m, w in wss.items()}
    torch.cuda.synchronize()
    assert not torch.equal(alone[ctr.INTERP_FP32], alone[ctr.INTERP_FAST])
    results = {m: [] for m in wss}
    errors = []
    barrier = threading.Barrier(len(wss))

    def worker(m):
        try:
            s = torch.cuda.Stream()
            barrier.wait()
            with torch.cuda.stream(s):
                for _ in range(6):
                    results[m].append(ctr.backproject(staged, g, ctr.F16, view_scale=vs,
                                                      scale=0.02, ws=wss[m], stream=s))
            s.synchronize()
        except Exception as e:  # reported by the main thread
            errors.append(e)

    threads = [threading.Thread(target=worker, args=(m,)) for m in wss]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=60)
    assert not errors, errors
    for m in wss:
        assert len(results[m]) == 6
        for r in results[m]:
            assert torch.equal(r, alone[m]), m
