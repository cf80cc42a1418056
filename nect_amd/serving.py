"""Continuous-batching export server for slice / sub-volume / volume queries.

The reference answers export queries one request at a time, one z-slice per
model call, with a host synchronisation per slice (nect/dynamic_export.py:
519-541, nect/static_export.py:106-141).  Here requests are served together:

  * every request is a voxel grid (torch.linspace coordinates per axis, a
    time t) visited in bricks of up to 4 x 4 x 4 voxels (``traversal``), so
    neighbouring voxels, which share hash-table entries, are evaluated
    together; the traversal is cut into chunks of ``chunk`` positions -- the
    scheduling quantum, independent of the request mix;
  * a request is admitted once it has arrived and its output buffer fits the
    HBM budget; the buffer lives on the device until the request completes;
  * each launch (one ``nect_query`` call) packs up to ``max_jobs`` chunks taken
    round-robin from the admitted requests in arrival order, so a full 512^3
    volume never blocks the slices that arrive behind it;
  * up to ``max_inflight`` launches are queued on the stream and completion is
    detected with events, so arrivals join while earlier work still runs.

A voxel's value depends only on its coordinates and the parameters (the query
kernel has no atomics or reductions), so each result is bit-identical whether
the request runs alone or in a mix (tests/test_serving_gpu.py).
"""
from __future__ import annotations

import time
from collections import deque
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from .grid import GridDesc
from .mlp import MlpDesc

# NectQueryJob (include/nect_amd.h)
JOB_DTYPE = np.dtype([("n", "<i4", 3), ("a", "<f4", 3), ("b", "<f4", 3), ("t", "<f4"), ("l_max", "<f4"),
                      ("span", "<f4"), ("lo", "<f4"), ("mu_dtype", "<i4"), ("rescale", "<i4"), ("brick", "<i4"),
                      ("axis", "<u8", 3), ("v0", "<i8"), ("count", "<i8"), ("out", "<u8")])
assert JOB_DTYPE.itemsize == 112

# NectRescale: how the network output becomes a voxel value, ((mu / l_max) * span) + lo
RESCALE_DIVIDE, RESCALE_RECIPROCAL, RESCALE_RECIPROCAL_F16 = 0, 1, 2


@dataclass
class QueryRequest:
    """A voxel grid: n (z, y, x) voxels on linspace(a[k], b[k], n[k]) per axis (torch's GPU rule) or on the
    explicit coordinates axes[k] (float32 GPU tensors of n[k] values), at time t.  The voxel value is
    ((mu / l_max) * span) + lo computed as ``rescale`` says (RESCALE_*), with mu first rounded to half when
    mu_half is set (the identity by default)."""

    n: tuple
    a: tuple
    b: tuple
    t: float = 0.0
    kind: str = "custom"
    arrival: float = 0.0
    rid: int = 0
    l_max: float = 1.0
    span: float = 1.0
    lo: float = 0.0
    mu_half: bool = False
    rescale: int = RESCALE_RECIPROCAL
    axes: tuple | None = None
    brick: int = 0
    slots: int = 0
    out: torch.Tensor | None = None
    host: torch.Tensor | None = None
    admitted: float | None = None
    done: float | None = None
    next_voxel: int = 0
    last_event: object = None

    @property
    def voxels(self) -> int:
        return int(self.n[0]) * int(self.n[1]) * int(self.n[2])

    @property
    def latency(self) -> float | None:
        return None if self.done is None else self.done - self.arrival


def traversal(n, bricks: bool = True) -> tuple[int, int]:
    """(brick code, traversal length) of an n = (z, y, x) grid: bricks of up to 4 voxels per axis (the
    largest power of two not above min(4, n_k)), which keep neighbouring voxels -- sharing hash-table
    entries -- together; the traversal covers the grid padded to whole bricks.  Row order (code 0) when
    ``bricks`` is off or every axis is a single voxel."""
    n = [int(k) for k in n]
    vox = n[0] * n[1] * n[2]
    if not bricks:
        return 0, vox
    lg = [min(2, k.bit_length() - 1) for k in n]
    code = (lg[0] << 8) | (lg[1] << 4) | lg[2]
    if code == 0:
        return 0, vox
    nb = [-(-k // (1 << l)) for k, l in zip(n, lg)]
    if nb[0] * nb[1] * nb[2] >= 1 << 31:
        return 0, vox
    return code, nb[0] * nb[1] * nb[2] << sum(lg)


def slice_request(axis: str, position: float, t: float, size: int = 512, **kw) -> QueryRequest:
    """An orthogonal size x size slice through the unit cube ("xy": fixed z, "xz": fixed y, "yz": fixed x)."""
    if axis == "xy":
        return QueryRequest((1, size, size), (position, 0.0, 0.0), (position, 1.0, 1.0), t, kind="slice", **kw)
    if axis == "xz":
        return QueryRequest((size, 1, size), (0.0, position, 0.0), (1.0, position, 1.0), t, kind="slice", **kw)
    if axis == "yz":
        return QueryRequest((size, size, 1), (0.0, 0.0, position), (1.0, 1.0, position), t, kind="slice", **kw)
    raise ValueError(axis)


def subvolume_request(origin, side: int, t: float, full: int = 512, **kw) -> QueryRequest:
    """An axis-aligned cube of side voxels of the full^3 export grid (endpoints included)."""
    a = tuple(o / (full - 1) for o in origin)
    b = tuple((o + side - 1) / (full - 1) for o in origin)
    return QueryRequest((side, side, side), a, b, t, kind="subvolume", **kw)


def volume_request(t: float, size: int = 512, **kw) -> QueryRequest:
    return QueryRequest((size, size, size), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), t, kind="volume", **kw)


def mixed_workload(rate: float, duration: float, seed: int = 0, size: int = 512) -> list[QueryRequest]:
    """Seeded Poisson arrivals: 70% orthogonal size^2 slices, 25% cubes of side U{size/8..size/2}
    (64..256 at size 512), 5% full size^3 volumes."""
    rng = np.random.default_rng(seed)
    reqs, t = [], 0.0
    while True:
        t += rng.exponential(1.0 / rate)
        if t > duration:
            break
        u = rng.random()
        tt = float(rng.random())
        if u < 0.70:
            r = slice_request(["xy", "xz", "yz"][int(rng.integers(3))], float(rng.random()), tt, size)
        elif u < 0.95:
            side = int(rng.integers(size // 8, size // 2 + 1))
            origin = tuple(int(rng.integers(0, size - side + 1)) for _ in range(3))
            r = subvolume_request(origin, side, tt, size)
        else:
            r = volume_request(tt, size)
        r.arrival = t
        r.rid = len(reqs)
        reqs.append(r)
    return reqs


class QueryServer:
    """Serves QueryRequests against one model held on the device (grid table + MLP weights in one dtype)."""

    def __init__(self, grid: GridDesc, mlp: MlpDesc, weights: torch.Tensor, table: torch.Tensor,
                 hbm_budget: int = 64 << 30, chunk: int = 1 << 18, max_jobs: int = 32, max_inflight: int = 3,
                 out_dtype: torch.dtype = torch.float16, bricks: bool = True, host_copy: bool = False,
                 timing: bool = False):
        if weights.dtype != table.dtype:
            raise ValueError("weights and table must share a dtype")
        if not table.is_cuda:
            raise ValueError("the model must be on the GPU")
        self.grid, self.mlp = grid, mlp
        self.weights, self.table = weights, table
        self.dtype = table.dtype
        self.device = table.device
        self.hbm_budget = int(hbm_budget)
        self.chunk = int(chunk)
        self.max_jobs = int(max_jobs)
        self.max_inflight = int(max_inflight)
        self.out_dtype = out_dtype
        self.bricks = bool(bricks)
        # host_copy: a request completes when its result is in pinned host memory (the reference's exports end
        # with .cpu()); the copies run on a second stream behind the launch that finishes the request
        self.host_copy = bool(host_copy)
        self._copy_stream = torch.cuda.Stream(self.device) if host_copy else None
        self._pinned: dict = {}
        # timing: start/end events around every nect_query launch (kernel_ms)
        self.timing = bool(timing)
        self._launch_events: list = []
        self.launch_voxels: list = []
        self.out_code = _lib.F16 if out_dtype == torch.float16 else _lib.F32
        self.esz = 2 if out_dtype == torch.float16 else 4
        self.in_use = 0
        slots = self.max_inflight + 2
        self._jobs_host = torch.empty((slots, max_jobs * JOB_DTYPE.itemsize), dtype=torch.uint8).pin_memory()
        self._pref_host = torch.empty((slots, max_jobs + 1), dtype=torch.int64).pin_memory()
        self._jobs_dev = torch.empty((slots, max_jobs * JOB_DTYPE.itemsize), dtype=torch.uint8, device=self.device)
        self._pref_dev = torch.empty((slots, max_jobs + 1), dtype=torch.int64, device=self.device)
        self._slot_events: list = [None] * slots
        self._slot = 0
        self.launches = 0

    # ------------------------------------------------------------- launches

    def launch(self, chunks: list) -> torch.cuda.Event:
        """Queues one nect_query launch over (request, first voxel, count) chunks; returns its event."""
        slot = self._slot
        self._slot = (slot + 1) % len(self._slot_events)
        if self._slot_events[slot] is not None:
            self._slot_events[slot].synchronize()  # the pinned staging slot is reused only after its copy ran
        jobs = np.zeros(len(chunks), JOB_DTYPE)
        pref = self._pref_host[slot].numpy()
        pref[0] = 0
        for k, (r, v0, cnt) in enumerate(chunks):
            jobs[k]["n"] = r.n
            jobs[k]["a"] = r.a
            jobs[k]["b"] = r.b
            jobs[k]["t"] = r.t
            jobs[k]["l_max"] = r.l_max
            jobs[k]["span"] = r.span
            jobs[k]["lo"] = r.lo
            jobs[k]["mu_dtype"] = _lib.F16 if r.mu_half else _lib.F32
            jobs[k]["rescale"] = r.rescale
            jobs[k]["brick"] = r.brick
            if r.axes is not None:
                for d in range(3):
                    ax = r.axes[d]
                    if ax is not None:
                        if ax.dtype != torch.float32 or not ax.is_cuda or ax.numel() != r.n[d] or not ax.is_contiguous():
                            raise ValueError("query axes must be contiguous float32 GPU tensors of n[k] values")
                        jobs[k]["axis"][d] = ax.data_ptr()
            jobs[k]["v0"] = v0
            jobs[k]["count"] = cnt
            jobs[k]["out"] = r.out.data_ptr()
            pref[k + 1] = pref[k] + (cnt + 31) // 32
        host = self._jobs_host[slot].numpy()
        host[: jobs.nbytes] = jobs.view(np.uint8)
        self._jobs_dev[slot].copy_(self._jobs_host[slot], non_blocking=True)
        self._pref_dev[slot].copy_(self._pref_host[slot], non_blocking=True)
        n_tiles = int(pref[len(chunks)])
        if self.timing:
            start = torch.cuda.Event(enable_timing=True)
            start.record()
        _lib.call("nect_query", self.grid.struct_ref(), self.mlp.struct_ref(), _lib.dtype_code(self.dtype),
                  self.weights.data_ptr(), self.table.data_ptr(), self._jobs_dev[slot].data_ptr(),
                  self._pref_dev[slot].data_ptr(), len(chunks), n_tiles, self.out_code, _lib.stream_of(self.device))
        if self.timing:
            end = torch.cuda.Event(enable_timing=True)
            end.record()
            self._launch_events.append((start, end))
            self.launch_voxels.append(sum(c[2] for c in chunks))
        ev = torch.cuda.Event()
        ev.record()
        self._slot_events[slot] = ev
        self.launches += 1
        return ev

    def kernel_ms(self) -> list:
        """(duration ms, traversal positions) of each timed launch so far (synchronises); clears the record."""
        torch.cuda.synchronize(self.device)
        out = [(a.elapsed_time(b), n) for (a, b), n in zip(self._launch_events, self.launch_voxels)]
        self._launch_events, self.launch_voxels = [], []
        return out

    def _to_host(self, finishing: list, after: torch.cuda.Event) -> torch.cuda.Event:
        """Queues the copies of finished results to pinned host buffers behind ``after``."""
        st = self._copy_stream
        st.wait_event(after)
        with torch.cuda.stream(st):
            for r in finishing:
                pool = self._pinned.setdefault(r.voxels, [])
                r.host = pool.pop() if pool else torch.empty(r.voxels, dtype=self.out_dtype, pin_memory=True)
                r.host.copy_(r.out, non_blocking=True)
            ev = torch.cuda.Event()
            ev.record(st)
        return ev

    def release_host(self, r: QueryRequest):
        """Returns a delivered result's pinned buffer to the pool."""
        if getattr(r, "host", None) is not None:
            self._pinned.setdefault(r.voxels, []).append(r.host)
            r.host = None

    def _alloc(self, r: QueryRequest):
        r.out = torch.empty(r.voxels, dtype=self.out_dtype, device=self.device)
        r.brick, r.slots = traversal(r.n, self.bricks)
        self.in_use += r.voxels * self.esz

    def run_alone(self, r: QueryRequest) -> torch.Tensor:
        """Serves a single request by itself (reference answer for the determinism check)."""
        self._alloc(r)
        v = 0
        while v < r.slots:
            chunks = []
            while v < r.slots and len(chunks) < self.max_jobs:
                cnt = min(self.chunk, r.slots - v)
                chunks.append((r, v, cnt))
                v += cnt
            self.launch(chunks)
        torch.cuda.synchronize(self.device)
        self.in_use -= r.voxels * self.esz
        return r.out.view(r.n)

    # ------------------------------------------------------------------ loop

    def serve(self, requests: list[QueryRequest], keep_outputs: bool = False) -> dict:
        """Replays requests at their arrival times (seconds from now) and serves them to completion."""
        pending = deque(sorted(requests, key=lambda r: r.arrival))
        active: list[QueryRequest] = []
        inflight: deque = deque()  # (event, [requests completing with it])
        done: list[QueryRequest] = []
        t0 = time.perf_counter()
        rr = 0
        busy = 0.0
        while pending or active or inflight:
            now = time.perf_counter() - t0
            # admission: arrived and fits the budget
            while pending and pending[0].arrival <= now:
                r = pending[0]
                need = r.voxels * self.esz
                if self.in_use + need > self.hbm_budget and (active or inflight):
                    break
                pending.popleft()
                self._alloc(r)
                r.admitted = now
                r.next_voxel = 0
                active.append(r)
            # completions
            while inflight and inflight[0][0].query():
                ev, finishing = inflight.popleft()
                now = time.perf_counter() - t0
                for r in finishing:
                    r.done = now
                    done.append(r)
                    if not keep_outputs:
                        self.in_use -= r.voxels * self.esz
                        r.out = None
                        self.release_host(r)
            # a new launch: chunks round-robin across the active requests
            if active and len(inflight) < self.max_inflight:
                chunks = []
                order = active[rr % len(active):] + active[: rr % len(active)]
                progressed = True
                while len(chunks) < self.max_jobs and progressed:
                    progressed = False
                    for r in order:
                        if r.next_voxel >= r.slots or len(chunks) >= self.max_jobs:
                            continue
                        cnt = min(self.chunk, r.slots - r.next_voxel)
                        chunks.append((r, r.next_voxel, cnt))
                        r.next_voxel += cnt
                        progressed = True
                rr += 1
                finishing = [r for r in active if r.next_voxel >= r.slots]
                active = [r for r in active if r.next_voxel < r.slots]
                ev = self.launch(chunks)
                if self.host_copy and finishing:
                    ev = self._to_host(finishing, ev)
                inflight.append((ev, finishing))
                continue
            if not inflight and not active and pending:
                wait = pending[0].arrival - (time.perf_counter() - t0)
                if wait > 0:
                    time.sleep(min(wait, 0.002))
            elif inflight:
                time.sleep(0.00005)
        wall = time.perf_counter() - t0
        return summarize(done, wall)


def summarize(done: list[QueryRequest], wall: float) -> dict:
    lat = {k: sorted(r.latency for r in done if r.kind == k) for k in ("slice", "subvolume", "volume")}
    allv = sorted(r.latency for r in done)

    def pct(v, q):
        if not v:
            return None
        return float(np.percentile(np.asarray(v), q) * 1e3)

    vox = sum(r.voxels for r in done)
    return {
        "requests": len(done), "wall_s": wall, "voxels": vox, "voxels_per_s": vox / wall if wall > 0 else 0.0,
        "p50_ms": pct(allv, 50), "p99_ms": pct(allv, 99),
        "slice_p50_ms": pct(lat["slice"], 50), "slice_p99_ms": pct(lat["slice"], 99),
        "subvolume_p50_ms": pct(lat["subvolume"], 50), "subvolume_p99_ms": pct(lat["subvolume"], 99),
        "volume_p50_ms": pct(lat["volume"], 50), "volume_p99_ms": pct(lat["volume"], 99),
        "counts": {k: len(v) for k, v in lat.items()},
    }
