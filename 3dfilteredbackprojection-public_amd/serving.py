"""Continuous-batching FDK reconstruction server on one GPU.

Requests arrive over time, each asking for one cone-beam FDK reconstruction
of one of three size classes:

small   256^3  from  360 views of  384^2
medium  512^3  from  720 views of  768^2
large   1024^3 from 1200 views of 1024^2

Each request carries its own seeded ellipsoid phantom: its projections are
the phantom's exact line integrals, generated on the device when the request
is admitted (ctr_project_ellipsoids, in the slot's stream ahead of the
pipeline; this stands in for the upload of a scan and counts toward the
request's latency).  Filtered projections are stored in FP16 with a per-view
scale, and the volume is FP32.

Design:
* A *slot* owns the device buffers of one reconstruction of a class: the
  input projections, the staged filtered projections, the per-view scale, the
  volume and a private ctr workspace.  It also owns its own HIP stream and a
  HIP graph of its whole pipeline (FDK weight + ramp filter into the staged
  layout, then backprojection with clipping), captured once and replayed for
  every request it serves.
* Small requests share *packed slots*: up to `pack` small reconstructions
  are filtered and then backprojected by ONE ctr_backproject_packed launch.
  There is one graph per fill count.
* Large reconstructions are backprojected in consecutive *view chunks*, one
  accumulating launch per chunk (the volume is read and written once per
  chunk).  A launch fills the whole GPU until it drains, so work of other
  slots waits for the launch in flight; with chunks it waits for one chunk
  instead of the whole reconstruction.
* Admission runs against an HBM budget.  A request is admitted as soon as a
  free slot of its class exists or a new one fits the budget; idle slots of
  other classes are evicted to make room.  Requests that arrive while others
  run are admitted as slots free up, and their work overlaps on the device
  through the slots' independent streams.  The queue is served FIFO, and a
  blocked head does not hold back requests of other classes.
* Latency is measured from arrival to the host observing the slot's
  completion event.  The result of a request does not depend on what else
  is in flight: every slot replays the same graph on its own buffers
  (tests/test_gpu_serving.py checks alone-vs-mixed equality).
"""
import collections
import dataclasses
import math
import statistics
import time
from dataclasses import dataclass, field

import numpy as np

from . import (F16, F32, MAX_PACKED, RAM_LAK, Workspace, backproject, backproject_packed,
               config_geometry, filter_projections, project_ellipsoids, random_ellipsoids,
               staged_layout)

CLASSES = {"small": "fdk_256_s", "medium": "fdk_512_s", "large": "fdk_1024_s"}


def fft_length_for(nu):
    """Zero padding to at least 2 nu - 1 samples (no circular wrap of the ramp)."""
    return 1 << int(2 * nu - 2).bit_length()


@dataclass
class Request:
    rid: int
    cls: str
    seed: int
    arrival: float                # seconds after the start of the run
    admitted: float = math.nan
    done: float = math.nan
    checksum: float = math.nan
    result: object = None         # volume copy when the server keeps results

    @property
    def latency(self):
        return self.done - self.arrival


def request_phantom(geom, seed):
    """The seeded ellipsoid phantom of one request: 30 ellipsoids per 512
    voxels of width (15 / 30 / 60 for the three classes), semi-axes 5 to
    120 voxels per 512, densities 0.1 .. 1.0 in units of 0.02 / mm."""
    n = max(4, int(30 * geom.X / 512))
    ell = random_ellipsoids(seed, n, (geom.X, geom.Y, geom.Z), geom.voxel,
                            (5, geom.X * 120 / 512))
    ell[:, 9] *= 0.02
    return ell


def make_projections(geom, seed, out=None, device=None, stream=None):
    """Exact projections of the request phantom of `seed`."""
    return project_ellipsoids(request_phantom(geom, seed), geom, out=out, device=device,
                              stream=stream)


def view_chunks(geom, chunks):
    """The views of `geom` split into `chunks` consecutive ranges: a list of
    (first view, geometry of the range).  Views i0 .. i0 + n - 1 of a scan of
    N views are themselves a scan of n views with rotation count rot n / N
    and offset offset + 2 pi rot i0 / N, since view i is at the angle
    2 pi rot i / N + offset (FBP3D.cpp:536-540)."""
    N = geom.n_views
    chunks = max(1, min(int(chunks), N))
    bounds = [N * c // chunks for c in range(chunks + 1)]
    return [(i0, dataclasses.replace(
                geom, n_views=i1 - i0, rotation_count=geom.rotation_count * (i1 - i0) / N,
                rotation_offset=geom.rotation_offset + 2 * math.pi * geom.rotation_count * i0 / N))
            for i0, i1 in zip(bounds, bounds[1:])]


class Slot:
    """Buffers, stream and captured graph for `count` reconstructions of one
    class (count > 1: a packed slot; a single one is backprojected in
    `chunks` view chunks)."""

    def __init__(self, cls, geom, dtype, count=1, use_graph=True, chunks=1):
        import torch
        self.cls, self.geom, self.dtype, self.count = cls, geom, dtype, count
        self.parts = view_chunks(geom, chunks) if count == 1 else [(0, geom)]
        self.fft = fft_length_for(geom.nu)
        self.scale = geom.fbp_scale(self.fft)
        dev = torch.device("cuda")
        self.stream = torch.cuda.Stream()
        self.ws = Workspace()
        g = geom
        _, _, sbytes = staged_layout(g.nu, g.nv, g.n_views, dtype)
        _, _, s32 = staged_layout(g.nu, g.nv, g.n_views, F32)
        self.proj = [torch.zeros((g.n_views, g.nv, g.nu), dtype=torch.float32, device=dev)
                     for _ in range(count)]
        self.staged = [torch.empty(sbytes, dtype=torch.uint8, device=dev) for _ in range(count)]
        self.vs = [torch.empty(g.n_views, dtype=torch.float32, device=dev) for _ in range(count)]
        self.vol = [torch.empty((g.Z, g.Y, g.X), dtype=torch.float32, device=dev)
                    for _ in range(count)]
        self.view_bytes = sbytes // g.n_views
        self.sums = torch.zeros(count, dtype=torch.float64, device=dev)
        self.bytes = count * (4 * g.n_views * g.nv * g.nu + sbytes + 4 * g.n_views +
                              4 * g.X * g.Y * g.Z) + (s32 if dtype != F32 else 0)
        self.busy = []            # requests in flight
        self.done_event = None
        self.last_used = 0.0
        self.graphs = {}
        with torch.cuda.stream(self.stream):
            self.ws.reserve(g.n_views * count, self.fft)
            for fill in range(1, count + 1):
                self._pipeline(fill)              # warm-up: sizes the workspace
            self.stream.synchronize()
            if use_graph:
                for fill in range(1, count + 1):
                    graph = torch.cuda.CUDAGraph()
                    with torch.cuda.graph(graph, stream=self.stream):
                        self._pipeline(fill)
                    self.graphs[fill] = graph
        self.stream.synchronize()

    def _pipeline(self, fill):
        g = self.geom
        vols = self.vol[:fill]
        for q in range(fill):
            filter_projections(self.proj[q], g, RAM_LAK, self.fft, self.dtype, fdk=True,
                               ws=self.ws, staged=self.staged[q], view_scale=self.vs[q])
        if self.count == 1:
            last = len(self.parts) - 1
            for c, (i0, gc) in enumerate(self.parts):
                i1 = i0 + gc.n_views
                backproject(self.staged[0][i0 * self.view_bytes:i1 * self.view_bytes], gc,
                            self.dtype, view_scale=self.vs[0][i0:i1], scale=self.scale,
                            volume=vols[0], accumulate=c > 0, clip=c == last, ws=self.ws)
        else:
            backproject_packed([(self.staged[q], g, self.vs[q], self.scale, vols[q], True)
                                for q in range(fill)], self.dtype, ws=self.ws)

    def launch(self, reqs, keep_results=False):
        import torch
        fill = len(reqs)
        with torch.cuda.stream(self.stream):
            for q, r in enumerate(reqs):
                make_projections(self.geom, r.seed, out=self.proj[q], stream=self.stream)
            if self.graphs:
                self.graphs[fill].replay()
            else:
                self._pipeline(fill)
            for q in range(fill):
                self.sums[q] = self.vol[q].sum(dtype=torch.float64)
            self.results = [v.clone() for v in self.vol[:fill]] if keep_results else None
            self.done_event = torch.cuda.Event()
            self.done_event.record(self.stream)
        self.busy = list(reqs)

    def poll(self):
        return self.done_event is not None and self.done_event.query()


class Server:
    """Admission, slot management and the serving loop."""

    def __init__(self, budget_bytes=200e9, dtype=F16, pack=4, max_slots=None,
                 use_graphs=True, keep_results=False, log=None, geometries=None, chunks=None):
        import torch
        self.device = torch.device("cuda")
        self.budget = float(budget_bytes)
        self.dtype = dtype
        self.pack = max(1, min(pack, MAX_PACKED))
        # one 1024^3 slot: a second one shares the GPU with the first and
        # delays every class (400 requests at 4 / s: p99 1240 ms against 983,
        # and 596 ms against 15 for 256^3)
        self.max_slots = max_slots or {"small": 2, "medium": 6, "large": 1}
        self.chunks = {"small": 1, "medium": 1, "large": 8, **(chunks or {})}
        self.use_graphs = use_graphs
        self.keep_results = keep_results
        self.geoms = geometries or {c: config_geometry(w) for c, w in CLASSES.items()}
        self.slots = []
        self.used = 0.0
        self.log = log or (lambda m: None)

    # --- slots -------------------------------------------------------------
    def _slot_count(self, cls):
        return self.pack if cls == "small" else 1

    def _estimate(self, cls):
        g = self.geoms[cls]
        count = self._slot_count(cls)
        _, _, sb = staged_layout(g.nu, g.nv, g.n_views, self.dtype)
        _, _, s32 = staged_layout(g.nu, g.nv, g.n_views, F32)
        return count * (4 * g.n_views * g.nv * g.nu + sb + 4 * g.n_views + 4 * g.X * g.Y * g.Z) \
            + (s32 if self.dtype != F32 else 0)

    def _free_slot(self, cls, now):
        for s in self.slots:
            if s.cls == cls and not s.busy:
                return s
        n_cls = sum(1 for s in self.slots if s.cls == cls)
        if n_cls >= self.max_slots[cls]:
            return None
        need = self._estimate(cls)
        if self.used + need > self.budget:
            # evict idle slots of other classes, least recently used first
            idle = sorted((s for s in self.slots if not s.busy and s.cls != cls),
                          key=lambda s: s.last_used)
            for s in idle:
                if self.used + need <= self.budget:
                    break
                self._drop(s)
            if self.used + need > self.budget:
                return None
        slot = self._new_slot(cls)
        self.slots.append(slot)
        self.used += slot.bytes
        self.log(f"slot {cls} x{slot.count}: {slot.bytes / 1e9:.2f} GB, in use "
                 f"{self.used / 1e9:.1f} / {self.budget / 1e9:.0f} GB")
        return slot

    def _new_slot(self, cls):
        return Slot(cls, self.geoms[cls], self.dtype, self._slot_count(cls), self.use_graphs,
                    self.chunks[cls])

    def _drop(self, slot):
        self.slots.remove(slot)
        self.used -= slot.bytes
        del slot

    def warm(self, classes):
        """Create the slots of the given classes ahead of the run, up to the
        per-class limits and the budget (graph capture synchronises the
        device, so it is kept out of every request's latency)."""
        made = []
        for c in classes:
            for _ in range(self.max_slots[c]):
                if self.used + self._estimate(c) > self.budget:
                    break
                s = self._new_slot(c)
                self.slots.append(s)
                self.used += s.bytes
                made.append(s)
        self.log(f"warm: {len(made)} slots, {self.used / 1e9:.1f} GB of "
                 f"{self.budget / 1e9:.0f} GB")
        return made

    # --- loop --------------------------------------------------------------
    def run(self, requests):
        """Serve `requests` (arrival times relative to the start) in real time."""
        pending = collections.deque(sorted(requests, key=lambda r: r.arrival))
        queue = []
        t0 = time.perf_counter()
        finished = 0
        total = len(pending)
        while finished < total:
            now = time.perf_counter() - t0
            while pending and pending[0].arrival <= now:
                queue.append(pending.popleft())
            progressed = False
            for s in self.slots:
                if s.busy and s.poll():
                    t = time.perf_counter() - t0
                    sums = s.sums[:len(s.busy)].cpu().tolist()
                    for q, r in enumerate(s.busy):
                        r.done = t
                        r.checksum = sums[q]
                        if s.results is not None:
                            r.result = s.results[q]
                    before = finished
                    finished += len(s.busy)
                    if self.log and finished // 100 > before // 100:
                        self.log(f"served {finished}/{total} at {t:.1f} s")
                    s.busy = []
                    s.last_used = t
                    progressed = True
            blocked = set()
            i = 0
            while i < len(queue):
                r = queue[i]
                if r.cls in blocked:
                    i += 1
                    continue
                slot = self._free_slot(r.cls, now)
                if slot is None:
                    blocked.add(r.cls)
                    i += 1
                    continue
                group = [r]
                if slot.count > 1:
                    for r2 in queue[i + 1:]:
                        if len(group) == slot.count:
                            break
                        if r2.cls == r.cls:
                            group.append(r2)
                for r2 in group:
                    queue.remove(r2)
                    r2.admitted = time.perf_counter() - t0
                slot.launch(group, self.keep_results)
                progressed = True
            if not progressed:
                nxt = pending[0].arrival - (time.perf_counter() - t0) if pending else 0.0
                time.sleep(min(max(nxt, 0.0), 0.0005) if nxt > 0 else 0.0002)
        return time.perf_counter() - t0


def poisson_arrivals(n, rate, seed, classes=("small", "medium", "large")):
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, size=n))
    cls = rng.integers(0, len(classes), size=n)
    return [Request(i, classes[c], int(rng.integers(0, 2**31)), float(t[i]))
            for i, c in enumerate(cls)]


def summarize(requests, makespan, geoms):
    lat = sorted(r.latency for r in requests)
    updates = sum(geoms[r.cls].updates for r in requests)
    by_cls = {}
    for c in {r.cls for r in requests}:
        lc = sorted(r.latency for r in requests if r.cls == c)
        by_cls[c] = {"n": len(lc), "p50_ms": 1e3 * lc[len(lc) // 2],
                     "p99_ms": 1e3 * lc[min(len(lc) - 1, int(0.99 * len(lc)))]}
    return {
        "requests": len(requests),
        "makespan_s": makespan,
        "p50_ms": 1e3 * statistics.median(lat),
        "p99_ms": 1e3 * lat[min(len(lat) - 1, int(math.ceil(0.99 * len(lat))) - 1)],
        "mean_ms": 1e3 * statistics.mean(lat),
        "aggregate_gups": updates / makespan / 1e9,
        "throughput_rps": len(requests) / makespan,
        "per_class": by_cls,
    }
