"""Serving loop: integration requests arrive over time, are admitted by HBM footprint as slots
and memory free up, keep their state resident on the device between steps, and run
concurrently (one context and HIP stream per request, work enqueued in chunks so kernels of
different requests overlap on the GPU). Reports p50/p99 completion latency and aggregate
cell-updates/s.

Requests run in the mixed form; a request whose mixed result is not finite is run again in
FP64 (fallback), and a request may ask for FP64 directly. Every kernel is deterministic and a
request's launches touch only its own buffers, so a request's result is the same alone and in
the mix, bit for bit.
"""
from __future__ import annotations

import collections
import ctypes
import time
from dataclasses import dataclass, field

import numpy as np

from . import _lib as L
from .ensemble import SurfaceState, draw_members, member_params
from .integration import Integration, footprint_bytes

GRID_CHOICES = [(21, 91, 181), (41, 181, 361), (61, 361, 721)]


@dataclass
class Request:
    rid: int
    arrival: float                 # seconds after the start of the run
    grid: tuple
    state: SurfaceState
    steps: int = 50
    precision: str = "mixed"
    seed: int = 0
    # filled in by the server
    admitted: float = -1.0
    finished: float = -1.0
    fallback: bool = False
    result: dict = field(default_factory=dict)

    @property
    def cells(self) -> int:
        return self.grid[0] * self.grid[1] * self.grid[2]

    @property
    def latency(self) -> float:
        return self.finished - self.arrival


def poisson_requests(n: int, rate_per_s: float = 4.0, seed: int = 99, steps: int = 50,
                     grids=GRID_CHOICES, precision: str = "mixed") -> list[Request]:
    """n requests with exponential inter-arrival times (mean 1/rate), grids drawn uniformly
    from `grids`, surface states as in the ensemble."""
    rng = np.random.default_rng(seed)
    states = draw_members(n, seed=seed + 1)
    t = 0.0
    out = []
    for q in range(n):
        t += float(rng.exponential(1.0 / rate_per_s))
        g = grids[int(rng.integers(len(grids)))]
        out.append(Request(q, t, tuple(g), states[q], steps, precision, seed=1000 + q))
    return out


def device_memory() -> tuple[int, int]:
    f, t = ctypes.c_int64(), ctypes.c_int64()
    L.check(L.lib().athad_device_memory(ctypes.byref(f), ctypes.byref(t)))
    return f.value, t.value


class Server:
    """Continuous-batching loop over per-request contexts.

    Finished contexts go back to a pool keyed by (grid, precision) and are reused for later
    requests of the same shape (new member parameters, fresh initial state), so the loop does
    no device allocation or free while requests run; pooled memory counts against the budget
    and idle pooled contexts are released when an admission needs the room. A request's result
    is the (sum, sum of squares) of each prognostic field, reduced on the device in a fixed
    order; keep_fields also downloads the fields."""

    def __init__(self, hbm_budget: int | None = None, max_slots: int = 4, chunk: int = 4,
                 keep_fields: bool = False):
        if hbm_budget is None:
            free, _ = device_memory()
            hbm_budget = int(free * 0.8)
        self.hbm_budget = hbm_budget
        self.max_slots = max_slots
        self.chunk = chunk
        self.keep_fields = keep_fields
        self.pool: dict = collections.defaultdict(list)
        self.used = 0                       # bytes held by active and pooled contexts
        self.peak_active = 0
        self.peak_bytes = 0
        self.reused = 0

    def close(self):
        for lst in self.pool.values():
            for ctx in lst:
                ctx.close()
        self.pool.clear()
        self.used = 0

    def _evict_for(self, need: int) -> bool:
        for key in list(self.pool):
            while self.pool[key] and self.used + need > self.hbm_budget:
                self.pool[key].pop().close()
                self.used -= footprint_bytes(key[0], key[1])
        return self.used + need <= self.hbm_budget

    def _admit(self, r: Request, now: float):
        """A slot for r, or None if the device memory is not there yet."""
        key = (r.grid, r.precision)
        if self.pool[key]:
            ctx = self.pool[key].pop()
            ctx.set_member_params(member_params(r.state))
            self.reused += 1
        else:
            need = footprint_bytes(r.grid, r.precision)
            if self.used + need > self.hbm_budget and not self._evict_for(need):
                return None
            ctx = Integration(r.grid, [member_params(r.state)], precision=r.precision)
            self.used += need
        ctx.init_state(r.seed)
        r.admitted = now
        return {"req": r, "ctx": ctx, "done": 0}

    def _finish(self, slot, now: float) -> bool:
        """Collects a finished request; False if it has to run again in FP64."""
        r, ctx = slot["req"], slot["ctx"]
        stats = {n: ctx.field_stats(n) for n in L.PROGNOSTIC}
        fields = {n: ctx.get(n) for n in L.PROGNOSTIC} if self.keep_fields else None
        self.pool[(r.grid, r.precision)].append(ctx)
        finite = all(np.isfinite(a) and np.isfinite(b) for a, b in stats.values())
        if not finite and r.precision == "mixed":
            r.precision = "fp64"
            r.fallback = True
            return False
        r.finished = now
        r.result = dict(stats)
        if fields is not None:
            r.result["fields"] = fields
        return True

    def run(self, requests: list[Request]) -> dict:
        pending = collections.deque(sorted(requests, key=lambda r: r.arrival))
        active = []
        t0 = time.perf_counter()
        while pending or active:
            now = time.perf_counter() - t0
            # admit arrived requests while a slot and device memory are free
            while pending and pending[0].arrival <= now and len(active) < self.max_slots:
                slot = self._admit(pending[0], now)
                if slot is None:
                    if not active:
                        raise RuntimeError(f"request {pending[0].rid} does not fit the budget {self.hbm_budget} B")
                    break
                pending.popleft()
                active.append(slot)
            self.peak_active = max(self.peak_active, len(active))
            self.peak_bytes = max(self.peak_bytes, self.used)
            # keep every admitted request's stream fed; collect the finished ones
            progressed = False
            for slot in list(active):
                if slot["ctx"].busy():
                    continue
                r = slot["req"]
                if slot["done"] >= r.steps:
                    active.remove(slot)
                    if not self._finish(slot, time.perf_counter() - t0):
                        pending.appendleft(r)          # FP64 fallback, ahead of later arrivals
                    progressed = True
                    continue
                n = min(self.chunk, r.steps - slot["done"])
                slot["ctx"].run(n, graph=True)
                slot["done"] += n
                progressed = True
            if not progressed:
                if not active and pending:
                    wait = pending[0].arrival - (time.perf_counter() - t0)
                    if wait > 0:
                        time.sleep(min(wait, 0.05))
                else:
                    time.sleep(2e-4)
        makespan = time.perf_counter() - t0
        lat = np.array([r.latency for r in requests])
        cells = sum(r.cells * r.steps for r in requests)
        return {"requests": len(requests), "makespan_s": makespan,
                "p50_latency_s": float(np.percentile(lat, 50)), "p99_latency_s": float(np.percentile(lat, 99)),
                "cells_per_s": cells / makespan, "peak_active": self.peak_active,
                "peak_hbm_bytes": self.peak_bytes, "contexts_reused": self.reused,
                "fallbacks": sum(r.fallback for r in requests)}
