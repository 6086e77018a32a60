"""Packed multi-volume backprojection and the continuous-batching server.

* ctr_backproject_packed gives, for every packed job, exactly what
  ctr_backproject gives for that job alone (mixed sizes and geometries).
* A request reconstructed alone and the same request served in a mix of
  concurrently admitted requests (different classes, packed small ones,
  several slots and streams in flight) give the same volume -- on toy sizes
  and at the serving workload's real sizes (1024^3 from 1200 views of 1024^2
  mixed with packed 256^3 and a 512^3), where one served volume is also
  checked against the FP64 oracle on a crop.
* A reconstruction served in view chunks equals the single launch up to
  FP32 rounding.
* Each request's projections are its own seeded phantom's exact line
  integrals (ctr_project_ellipsoids, checked against the oracle's analytic
  projector).
"""
import numpy as np
import pytest

from conftest import error_metrics, format_metrics, rel_max

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
    mk = lambda: serving.Server(budget_bytes=4e9, pack=3, keep_results=True,
                                geometries=geoms,
                                max_slots={"small": 2, "medium": 2, "large": 2})
    mix = [serving.Request(i, c, s, 0.0) for i, (c, s) in enumerate(
        [("large", 0), ("small", 1), ("small", 2), ("medium", 0), ("small", 0),
         ("large", 1), ("small", 1), ("medium", 2), ("small", 2), ("medium", 1)])]
    srv = mk()
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


@pytest.mark.parametrize("chunks", [2, 5])
def test_view_chunks_match_one_launch(ctr, dev, chunks):
    """A reconstruction backprojected in view chunks (one accumulating launch
    per chunk, clipped after the last, replayed from the slot's graph) equals
    the single launch up to the FP32 rounding of its partial sums."""
    import torch
    from ctrecon import serving
    g = _geoms(ctr)["large"]
    one = serving.Slot("large", g, ctr.F16)
    many = serving.Slot("large", g, ctr.F16, chunks=chunks)
    assert len(many.parts) == chunks
    out = []
    for slot in (one, many):
        for seed in (3, 4):
            slot.launch([serving.Request(seed, "large", seed, 0.0)], keep_results=True)
            torch.cuda.synchronize()
            out.append(slot.results[0].cpu().numpy())
    for a, b in zip(out[:2], out[2:]):
        assert np.abs(a).max() > 0
        assert rel_max(b, a) < 2e-6, rel_max(b, a)


def test_server_respects_budget(ctr, dev):
    from ctrecon import serving
    geoms = _geoms(ctr)
    srv = serving.Server(budget_bytes=0.05e9, pack=2, geometries=geoms)
    est = {c: srv._estimate(c) for c in geoms}
    srv.warm(["large", "small"])
    assert srv.used <= 0.05e9
    assert all(s.bytes <= est[s.cls] for s in srv.slots)


@pytest.mark.parametrize("kind", ["cone", "parallel"])
def test_project_ellipsoids_vs_oracle(ctr, orc, dev, kind):
    """The analytic projector against the oracle's FP64 line integrals (both
    layouts, more ellipsoids than one launch carries, accumulate)."""
    import torch
    ell = ctr.random_ellipsoids(8, 70, (64, 48, 64), 1.0, (2, 25))
    if kind == "cone":
        g = ctr.Geometry(ctr.CONE, nu=96, nv=72, n_views=20, X=64, Y=48, Z=64,
                         detector_width=200.0, volume_width=64.0, source_distance=150.0,
                         detector_distance=120.0, rotation_offset=0.3)
        ref = orc.analytic_cone(ell, g.source_distance, g.detector_distance, g.detector_width,
                                g.nu, g.nv, g.n_views, 1.0, 0.3)
    else:
        g = ctr.Geometry(ctr.PARALLEL, nu=96, nv=72, n_views=20, X=64, Y=48, Z=64,
                         detector_width=100.0, volume_width=64.0, rotation_count=0.5,
                         rotation_offset=0.3)
        ref = orc.analytic_parallel(ell, g.detector_width, g.nu, g.nv, g.n_views, 0.5, 0.3)
    got = ctr.project_ellipsoids(ell, g).cpu().numpy()
    assert ref.max() > 1.0
    # FP64 chords, rounded to FP32 once
    assert np.abs(got - ref).max() <= 2e-7 * np.abs(ref).max()
    sino = ctr.project_ellipsoids(ell, g, layout=ctr.SINOGRAMS).cpu().numpy()
    assert np.array_equal(sino.transpose(1, 0, 2), got)
    twice = ctr.project_ellipsoids(ell, g, out=torch.from_numpy(got.copy()).to(dev),
                                   accumulate=True).cpu().numpy()
    assert np.allclose(twice, 2 * got, rtol=1e-6, atol=0)


def test_serving_at_real_sizes(ctr, orc, dev):
    """The serving workload's classes at their real sizes: one 1024^3 request
    (1200 views of 1024^2) admitted with five 256^3 requests (360 views of
    384^2; packed three to a launch) and one 512^3 (720 views of 768^2), FP16
    storage.  Every result equals the same request served alone, and the
    1024^3 volume matches the FP64 oracle pipeline (FDK weight, reference
    ramp filter, voxel loop) of the same projections on a 16^3 crop."""
    import torch
    from ctrecon import serving
    mix = [serving.Request(i, c, s, 0.0) for i, (c, s) in enumerate(
        [("large", 11), ("small", 1), ("small", 2), ("medium", 3), ("small", 4),
         ("small", 5), ("small", 6)])]
    srv = serving.Server(budget_bytes=200e9, pack=3, keep_results=True,
                         max_slots={"small": 2, "medium": 1, "large": 1})
    srv.warm(["small", "medium", "large"])
    print("[serving sizes] slots warm", flush=True)
    srv.run(mix)
    print("[serving sizes] mix served", flush=True)
    for r in mix:
        alone = serving.Request(100 + r.rid, r.cls, r.seed, 0.0)
        srv.run([alone])
        assert torch.equal(alone.result, r.result), (r.cls, r.seed)
        assert alone.checksum == r.checksum
    print("[serving sizes] alone == mixed", flush=True)
    large = mix[0]
    g = srv.geoms["large"]
    assert (g.X, g.n_views, g.nu) == (1024, 1200, 1024)
    box = ((504, 520), (504, 520), (500, 516))
    # the crop projects within +-12 detector rows of the centre: the oracle
    # needs only a band of rows, which the FDK weight and the row filter
    # treat exactly as the full panel (same pitch, same centre)
    h = 64
    r0 = g.nv // 2 - h // 2
    proj = serving.make_projections(g, large.seed, device=dev)
    band = proj[:, r0:r0 + h, :].cpu().numpy()
    del proj
    w = orc.fdk_weight(band, g.source_distance, g.detector_distance, g.detector_width)
    slot = next(s for s in srv.slots if s.cls == "large")
    q = orc.filter_rows(w.reshape(-1, g.nu), orc.RAM_LAK, slot.fft).reshape(w.shape)
    ref = orc.backproject_cone_region(q.astype(np.float32), g.source_distance,
                                      g.detector_distance, g.detector_width, g.volume_width,
                                      1.0, 0.0, slot.scale, (g.X, g.Y, g.Z), box)
    ref = np.maximum(ref, 0.0)  # the served pipeline clips negatives
    (i0, i1), (j0, j1), (k0, k1) = box
    got = large.result[k0:k1, j0:j1, i0:i1].cpu().numpy()
    m = error_metrics(got, ref)
    print(f"[serving sizes] 1024^3 crop vs oracle (FP16 store): {format_metrics(m)}", flush=True)
    assert np.abs(ref).max() > 0
    assert m["rel_rmse"] < 1e-3
