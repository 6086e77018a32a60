"""ctr_backproject_noalloc, the entry point that allocates nothing, against
ctr_backproject through the C ABI: bit for bit on two calls in a row, the
second smaller (its scratch array keeps the first call's contents), for the
LDS-staged and the direct-tap kernels; device memory does not change across
a call; jobs whose tables exceed the fixed scratch are refused before any
launch."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _geom(ctr, **kw):
    base = dict(kind=ctr.CONE, nu=96, nv=80, n_views=120, X=96, Y=64, Z=80,
                detector_width=180.0, volume_width=96.0, source_distance=220.0,
                detector_distance=160.0, rotation_offset=0.2)
    base.update(kw)
    return ctr.Geometry(**base)


def _noalloc(ctr, staged, g, dt, vs, scale, volume, accumulate=0, clip=0):
    return ctr.lib().ctr_backproject_noalloc(
        ctr.stream_handle(), g.kind, ctr.ptr(staged), dt, ctr.ptr(vs), g.nu, g.nv, g.n_views,
        g.source_distance, g.detector_distance, g.detector_width, g.volume_width,
        g.rotation_count, g.rotation_offset, ctr.ptr(volume), g.X, g.Y, g.Z,
        ctypes.c_float(scale), accumulate, clip)


def test_noalloc_matches_backproject_twice(ctr, dev):
    import torch
    cases = [(_geom(ctr), ctr.F16, 0.02, 0, 0),
             (_geom(ctr, nu=64, nv=48, n_views=90, X=40, Y=32, Z=36, detector_width=128.0,
                    volume_width=40.0, source_distance=150.0, detector_distance=100.0),
              ctr.F32, 0.01, 1, 1),
             (_geom(ctr, nu=70, nv=8, n_views=60, X=48, Y=6, Z=44, detector_width=140.0,
                    volume_width=48.0, kind=ctr.PARALLEL, rotation_count=0.5),
              ctr.BF16, 0.03, 0, 1)]
    for q, (g, dt, scale, acc, clip) in enumerate(cases):
        gen = torch.Generator(device=dev)
        gen.manual_seed(500 + q)
        p = torch.randn((g.n_views, g.nv, g.nu), generator=gen, device=dev)
        staged, vs = ctr.stage_projections(p, g, dt)
        base = torch.randn((g.Z, g.Y, g.X), generator=gen, device=dev)
        ref = ctr.backproject(staged, g, dt, view_scale=vs, scale=scale, volume=base.clone(),
                              accumulate=bool(acc), clip=bool(clip))
        out = base.clone()
        torch.cuda.synchronize()
        free_before = torch.cuda.mem_get_info()[0]
        rc = _noalloc(ctr, staged, g, dt, vs, scale, out, acc, clip)
        assert rc == 0, ctr.last_error()
        torch.cuda.synchronize()
        if q > 0:  # the first call on a device may load the code object
            assert torch.cuda.mem_get_info()[0] == free_before
        assert torch.equal(out, ref), (q, g.X, g.Y, g.Z, dt)


def test_noalloc_refuses_jobs_beyond_its_scratch(ctr, dev):
    import torch
    # 1440^2 columns x 1440 views: 186 MB of footprint tables > 128 MiB
    g = _geom(ctr, nu=64, nv=64, n_views=1440, X=1440, Y=64, Z=1440, detector_width=600.0,
              volume_width=360.0, source_distance=500.0, detector_distance=500.0)
    t = torch.zeros(16, device=dev)
    rc = _noalloc(ctr, t, g, ctr.F32, None, 1.0, t)
    assert rc == -6
    assert "scratch" in ctr.last_error()
