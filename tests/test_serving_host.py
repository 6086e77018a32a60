"""Host logic of the serving loop, no GPU: the Poisson arrival stream, the
per-request phantoms and the latency summary."""
import math

import numpy as np
import pytest


@pytest.fixture(scope="module")
def serving(ctr):
    from ctrecon import serving
    return serving


def test_poisson_arrivals_reproducible_and_at_rate(serving):
    a = serving.poisson_arrivals(4000, 4.0, seed=7)
    b = serving.poisson_arrivals(4000, 4.0, seed=7)
    assert [(r.cls, r.seed, r.arrival) for r in a] == [(r.cls, r.seed, r.arrival) for r in b]
    t = np.array([r.arrival for r in a])
    assert np.all(np.diff(t) > 0)
    assert 4000 / t[-1] == pytest.approx(4.0, rel=0.05)          # mean rate
    gaps = np.diff(t)
    assert np.std(gaps) / np.mean(gaps) == pytest.approx(1.0, rel=0.06)  # exponential
    counts = {c: sum(r.cls == c for r in a) for c in ("small", "medium", "large")}
    assert all(abs(n - 4000 / 3) < 120 for n in counts.values())
    assert len({r.seed for r in a}) > 3990                        # a phantom per request


def test_request_phantoms_per_class(ctr, serving):
    for name, n_ell in (("fdk_256_s", 15), ("fdk_512_s", 30), ("fdk_1024_s", 60)):
        g = ctr.config_geometry(name)
        e = serving.request_phantom(g, 123)
        assert e.shape == (n_ell, 10)
        assert np.array_equal(e, serving.request_phantom(g, 123))
        assert not np.array_equal(e, serving.request_phantom(g, 124))
        half = 0.5 * g.X * g.voxel
        assert np.all(np.abs(e[:, :3]) <= 0.4 * 2 * half + 1e-9)   # inner 80% of the field
        semi = e[:, 3:6] / g.voxel
        assert semi.min() >= 5 - 1e-9 and semi.max() <= g.X * 120 / 512 + 1e-9
        assert np.all((e[:, 9] >= 0.1 * 0.02) & (e[:, 9] <= 1.0 * 0.02))


def test_summary_percentiles(ctr, serving):
    g = {c: ctr.config_geometry(n) for c, n in serving.CLASSES.items()}
    reqs = []
    for i in range(200):
        r = serving.Request(i, "small" if i % 2 else "large", i, arrival=0.01 * i)
        r.done = r.arrival + 0.001 * (i + 1)      # latencies 1 .. 200 ms
        reqs.append(r)
    s = serving.summarize(reqs, 2.0, g)
    assert s["p50_ms"] == pytest.approx(100.5, abs=1e-6)
    assert s["p99_ms"] == pytest.approx(198.0, abs=1e-6)
    assert s["throughput_rps"] == pytest.approx(100.0)
    updates = 100 * g["small"].updates + 100 * g["large"].updates
    assert s["aggregate_gups"] == pytest.approx(updates / 2.0 / 1e9)
    assert s["per_class"]["small"]["n"] == 100
    assert math.isclose(s["mean_ms"], 100.5)


@pytest.mark.parametrize("chunks", [1, 3, 7, 120])
def test_view_chunks_cover_the_scan(ctr, serving, chunks):
    """Each chunk is the scan restricted to consecutive views: together they
    hold every view once, and chunk view m of a chunk starting at i0 has the
    angle of scan view i0 + m (angle 2 pi rot i / N + offset)."""
    g = ctr.Geometry(ctr.CONE, nu=64, nv=48, n_views=120, X=40, Y=36, Z=40,
                     detector_width=120.0, volume_width=40.0, source_distance=150.0,
                     detector_distance=100.0, rotation_count=0.75, rotation_offset=0.3)
    parts = serving.view_chunks(g, chunks)
    assert len(parts) == chunks
    angle = lambda geo, i: 2 * math.pi * geo.rotation_count * i / geo.n_views + geo.rotation_offset
    seen = []
    for i0, gc in parts:
        assert gc.n_views >= 1 and (gc.nu, gc.nv, gc.X, gc.kind) == (g.nu, g.nv, g.X, g.kind)
        for m in range(gc.n_views):
            assert angle(gc, m) == pytest.approx(angle(g, i0 + m), abs=1e-12)
        seen += range(i0, i0 + gc.n_views)
    assert seen == list(range(g.n_views))
    assert serving.view_chunks(g, 1)[0][1] == g
