"""Batched ensemble: many integrations with their own surface state and composition,
advanced together in the same launches (one context, blockIdx.z = member), the 4-iteration
cycle captured once as a hipGraph and replayed.

Members are drawn from a seed uniformly over surface pressure 40-100 bar, sea temperature
503-523 K kept below the boiling point at that pressure, dry-air H2O 0.4-2 % and CO2 89-95 %
by mole (N2 the remainder), and each gets the parameters test/cond_column_selftest.cpp:86-145
derives for the reference's own configuration (saturated sea, CO2 : N2 at the dry ratio,
r_air from the surface state), so every member integrates its own moist-adiabat column.
"""
from __future__ import annotations

import ctypes
import time
from dataclasses import dataclass

import numpy as np

from . import _lib as L
from .integration import Integration


@dataclass
class SurfaceState:
    p_surf_bar: float
    t_sea: float
    x_h2o_dry: float
    x_co2_dry: float


def member_params(s: SurfaceState) -> L.MemberParams:
    out = L.MemberParams()
    L.check(L.lib().athad_member_from_surface(s.p_surf_bar, s.t_sea, s.x_h2o_dry, s.x_co2_dry,
                                              ctypes.byref(out)))
    return out


def dew_point(e_hpa: float) -> float:
    """Boiling temperature [K] of water at vapour pressure e [hPa]."""
    return float(L.lib().athad_dew_point(e_hpa))


def draw_members(n: int, seed: int = 2024, p_bar=(40.0, 100.0), t_sea=(503.0, 523.0),
                 x_h2o_dry=(0.004, 0.02), x_co2_dry=(0.89, 0.95)) -> list[SurfaceState]:
    """n surface states from a seed; a sea temperature at or above the boiling point at its
    pressure is drawn again."""
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < n:
        p = rng.uniform(*p_bar)
        t = rng.uniform(*t_sea)
        xh = rng.uniform(*x_h2o_dry)
        xc = rng.uniform(*x_co2_dry)
        if t >= dew_point(p * 1000.0):
            continue
        out.append(SurfaceState(p, t, xh, xc))
    return out


class Ensemble(Integration):
    """An Integration over several members (grids may differ)."""

    def __init__(self, grids, states: list[SurfaceState], precision: str = "fp64", order: str = "jacobi"):
        if isinstance(grids, tuple) and len(grids) == 3 and isinstance(grids[0], int):
            grids = [grids] * len(states)
        if len(grids) != len(states):
            raise ValueError("one grid per member")
        self.states = list(states)
        super().__init__(list(grids), [member_params(s) for s in states], precision=precision, order=order)


def run_ensemble(n_members=32, grid=(41, 91, 181), steps=50, warmup=4, precision="fp64", seed=2024,
                 order="jacobi"):
    """Times `steps` iterations of an n-member ensemble; returns aggregate cell-updates/s. In the
    lex order the members share every hyperplane launch."""
    states = draw_members(n_members, seed)
    cells = n_members * grid[0] * grid[1] * grid[2]
    with Ensemble(grid, states, precision=precision, order=order) as ens:
        ens.init_state(seed)
        ens.run(warmup, graph=True)
        ens.sync()
        t0 = time.perf_counter()
        ens.run(steps, graph=True)
        ens.sync()
        dt = time.perf_counter() - t0
        finite = all(np.isfinite(ens.get("t", m)).all() for m in range(n_members))
    return {"members": n_members, "grid": list(grid), "steps": steps, "precision": precision, "order": order,
            "cells_per_s": cells * steps / dt, "ms_per_step": dt / steps * 1e3, "finite": finite}
