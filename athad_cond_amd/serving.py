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
    """Continuous-batching loop over per-request contexts."""

    def __init__(self, hbm_budget: int | None = None, max_slots: int = 8, chunk: int = 4,
                 keep_fields: bool = False):
        if hbm_budget is None:
            free, _ = device_memory()
            hbm_budget = int(free * 0.8)
        self.hbm_budget = hbm_budget
        self.max_slots = max_slots
        self.chunk = chunk
        self.keep_fields = keep_fields
        self.peak_active = 0
        self.peak_bytes = 0

    def _start(self, r: Request, now: float):
        ctx = Integration(r.grid, [member_params(r.state)], precision=r.precision)
        ctx.init_state(r.seed)
        r.admitted = now
        return {"req": r, "ctx": ctx, "done": 0, "bytes": footprint_bytes(r.grid, r.precision)}

    def _finish(self, slot, now: float) -> bool:
        """Collects a finished request; False if it has to run again in FP64."""
        r, ctx = slot["req"], slot["ctx"]
        fields = {n: ctx.get(n) for n in L.PROGNOSTIC}
        ctx.close()
        finite = all(np.isfinite(v).all() for v in fields.values())
        if not finite and r.precision == "mixed":
            r.precision = "fp64"
            r.fallback = True
            return False
        r.finished = now
        r.result = {n: (float(v.sum()), float(np.sqrt((v * v).sum()))) for n, v in fields.items()}
        if self.keep_fields:
            r.result["fields"] = fields
        return True

    def run(self, requests: list[Request]) -> dict:
        pending = collections.deque(sorted(requests, key=lambda r: r.arrival))
        active = []
        used = 0
        t0 = time.perf_counter()
        while pending or active:
            now = time.perf_counter() - t0
            # admit arrived requests while a slot and device memory are free
            while pending and pending[0].arrival <= now and len(active) < self.max_slots:
                need = footprint_bytes(pending[0].grid, pending[0].precision)
                if used + need > self.hbm_budget:
                    if not active:
                        raise RuntimeError(f"request {pending[0].rid} needs {need} B, budget {self.hbm_budget} B")
                    break
                slot = self._start(pending.popleft(), now)
                used += slot["bytes"]
                active.append(slot)
            self.peak_active = max(self.peak_active, len(active))
            self.peak_bytes = max(self.peak_bytes, used)
            # keep every admitted request's stream fed; collect the finished ones
            progressed = False
            for slot in list(active):
                if slot["ctx"].busy():
                    continue
                r = slot["req"]
                if slot["done"] >= r.steps:
                    now = time.perf_counter() - t0
                    used -= slot["bytes"]
                    active.remove(slot)
                    if not self._finish(slot, now):
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
                "peak_hbm_bytes": self.peak_bytes, "fallbacks": sum(r.fallback for r in requests)}
