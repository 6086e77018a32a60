"""Packed multi-volume backprojection and the continuous-batching server.

* ctr_backproject_packed gives, for every packed job, exactly what
  ctr_backproject gives for that job alone (mixed sizes and geometries).
* A request reconstructed alone and the same request served in a mix of
  concurrently admitted requests (different classes, packed small ones,
  several slots and streams in flight) give the same volume.
"""
import numpy as np
import pytest

from conftest import rel_max

pytestmark = pytest.mark.gpu


def _geoms(ctr):
    return {
        "small": ctr.Geometry(ctr.CONE, nu=64, nv=48, n_views=60, X=40, Y=36, Z=40,
                              detector_width=120.0, volume_width=40.0, source_distance=150.0,
                              detector_distance=100.0),
        "medium": ctr.Geometry(ctr.CONE, nu=96, nv=80, n_views=90, X=64, Y=64, Z=64,
                               detector_width=180.0, volume_width=64.0, source_distance=200.0,
                               detector_distance=150.0),
        "large": ctr.Geometry(ctr.CONE, nu=128, nv=128, n_views=120, X=96, Y=100, Z=96,
                              detector_width=240.0, volume_width=96.0, source_distance=250.0,
                              detector_distance=200.0, rotation_offset=0.2),
    }


def test_packed_matches_single(ctr, dev):
    import torch
    gs = list(_geoms(ctr).values()) + [
        ctr.Geometry(ctr.PARALLEL, nu=50, nv=40, n_views=45, X=32, Y=33, Z=30,
                     detector_width=60.0, volume_width=40.0, rotation_count=0.5)]
    items, singles = [], []
    for q, g in enumerate(gs):
        gen = torch.Generator(device=dev)
        gen.manual_seed(q)
        p = torch.randn((g.n_views, g.nv, g.nu), generator=gen, device=dev)
        staged, vs = ctr.stage_projections(p, g, ctr.F16)
        vol = torch.empty((g.Z, g.Y, g.X), device=dev)
        items.append((staged, g, vs, 0.01 * (q + 1), vol, q % 2 == 0))
        singles.append(ctr.backproject(staged, g, ctr.F16, view_scale=vs, scale=0.01 * (q + 1),
                                       clip=q % 2 == 0))
    ctr.backproject_packed(items, ctr.F16)
    torch.cuda.synchronize()
    for it, ref in zip(items, singles):
        assert torch.equal(it[4], ref)


def test_alone_equals_mixed(ctr, dev):
    import torch
    from ctrecon import serving
    geoms = _geoms(ctr)
    mk = lambda: serving.Server(budget_bytes=4e9, pack=3, keep_results=True, pool=3,
                                geometries=geoms,
                                max_slots={"small": 2, "medium": 2, "large": 2})
    mix = [serving.Request(i, c, s, 0.0) for i, (c, s) in enumerate(
        [("large", 0), ("small", 1), ("small", 2), ("medium", 0), ("small", 0),
         ("large", 1), ("small", 1), ("medium", 2), ("small", 2), ("medium", 1)])]
    srv = mk()
    srv.prepare_sources(["small", "medium", "large"])
    srv.warm(["small", "medium", "large"])
    makespan = srv.run(mix)
    assert all(np.isfinite(r.latency) and r.latency >= 0 for r in mix)
    for r in mix:
        alone = serving.Request(100 + r.rid, r.cls, r.seed, 0.0)
        srv.run([alone])
        a, m = alone.result, r.result
        assert a is not None and m is not None
        assert torch.equal(a, m), (r.cls, r.seed, rel_max(a.cpu().numpy(), m.cpu().numpy()))
        assert alone.checksum == r.checksum
    summary = serving.summarize(mix, makespan, geoms)
    assert summary["requests"] == len(mix) and summary["aggregate_gups"] > 0


def test_server_respects_budget(ctr, dev):
    from ctrecon import serving
    geoms = _geoms(ctr)
    srv = serving.Server(budget_bytes=0.05e9, pack=2, pool=1, geometries=geoms)
    est = {c: srv._estimate(c) for c in geoms}
    srv.warm(["large", "small"])
    assert srv.used <= 0.05e9
    assert all(s.bytes <= est[s.cls] for s in srv.slots)
