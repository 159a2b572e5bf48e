"""GPU moist convection (MoistConvection::run, MoistConvection.h:106-140) against the CPU oracle.

The oracle is pinned to the reference's compiled scheme by
tests/test_oracle_golden.py::test_moist_convection_twice on the same kind of input: a spun-up
state with the condensate that oracle.plant_microphysics plants, and a seeded per-column
evaporation diagnostic, so that updrafts, downdrafts, parcel condensation and convective
precipitation are all live. Two calls in a row from zeroed arrays: the second reads what the
first left in the arrays the scheme does not clear. Tolerance 1e-11 relative L2 per array in
FP64 (the device's exp / pow / cbrt differ from glibc's by an ulp and a flipped threshold moves
one column; measured values are printed).
"""
import numpy as np
import pytest

import athad_cond_amd as A
from athad_cond_amd import _lib as L

pytestmark = pytest.mark.gpu

ARRAYS = L.CONV_FIELDS[:-1]


def rel(a, b):
    den = np.sqrt(np.sum(b * b))
    num = np.sqrt(np.sum((a - b) ** 2))
    return num / den if den > 0 else num


def planted(oracle, grid, seed, spin=2):
    ref = oracle.OracleState(*grid, seed=seed)
    ref.run(spin, "jacobi")
    snap = ref.snapshot()
    oracle.plant_microphysics(snap, {k: np.zeros(grid) for k in oracle.MP_FIELDS}, seed + 1)
    ref.load(snap)
    evap = np.random.default_rng(seed).uniform(-1.0, 6.0, grid[1:])
    ref.evap()[...] = evap
    return ref, snap, evap


def errors(integ, ref, member=0):
    return {n: rel(integ.get_conv(n, member=member), ref.mc(n)) for n in ARRAYS}


@pytest.mark.parametrize("grid", [(41, 181, 361), (61, 91, 73)])
def test_two_calls_match_oracle(oracle, grid):
    print(f"case moist convection fp64 {grid}", flush=True)
    ref, snap, evap = planted(oracle, grid, 21)
    with A.Integration(grid) as integ:
        integ.load(snap)
        integ.set_conv("Evaporation", evap)
        for call in (1, 2):
            integ.moist_convection()
            ref.moist_convection()
            e = errors(integ, ref)
            print(call, {k: f"{v:.1e}" for k, v in e.items() if v > 0})
            bad = {k: v for k, v in e.items() if v > 1e-11}
            assert not bad, (call, bad)
            # computeCAPE: each level from the same (last) column as the oracle's, zero where no
            # column reached it
            cape, cref = integ.conv_cape(), ref.cape()
            print(call, "cape", f"{rel(cape, cref):.1e}", int(np.count_nonzero(cref)), "levels set")
            assert np.count_nonzero(cref) > 0
            assert np.array_equal(cape == 0.0, cref == 0.0)
            assert rel(cape, cref) <= 1e-11
        # the input makes the scheme do something
        for n in ("M_u", "M_d", "q_c_u", "c_u", "e_d", "MC_t", "MC_q", "MC_v", "MC_w"):
            assert np.any(ref.mc(n) != 0.0), n
        assert np.all(np.isfinite(integ.get_conv("MC_t")))


def test_members_are_independent(oracle):
    """Two members with different inputs in one call: each matches its own oracle run."""
    grid = (41, 91, 73)
    print(f"case moist convection members {grid}", flush=True)
    refs = [planted(oracle, grid, s) for s in (5, 9)]
    with A.Integration([grid, grid]) as integ:
        for m, (ref, snap, evap) in enumerate(refs):
            integ.load(snap, member=m)
            integ.set_conv("Evaporation", evap, member=m)
        integ.moist_convection()
        for m, (ref, _, _) in enumerate(refs):
            ref.moist_convection()
            e = errors(integ, ref, member=m)
            bad = {k: v for k, v in e.items() if v > 1e-11}
            assert not bad, (m, bad)
            cape, cref = integ.conv_cape(member=m), ref.cape()
            assert np.array_equal(cape == 0.0, cref == 0.0), m
            assert rel(cape, cref) <= 1e-11, m


def test_mixed_precision_fields(oracle):
    """FP32 field storage: the scheme reads the rounded fields and computes in FP64; against
    the oracle run on the same rounded fields."""
    grid = (41, 91, 73)
    print(f"case moist convection mixed {grid}", flush=True)
    ref, snap, evap = planted(oracle, grid, 13)
    with A.Integration(grid, precision="mixed") as integ:
        integ.load(snap)
        integ.set_conv("Evaporation", evap)
        rounded = {n: integ.get(n) for n in snap}
        ref.load(rounded)
        integ.moist_convection()
        ref.moist_convection()
        e = errors(integ, ref)
        print({k: f"{v:.1e}" for k, v in e.items() if v > 0})
        bad = {k: v for k, v in e.items() if v > 1e-11}
        assert not bad, bad
        cape, cref = integ.conv_cape(), ref.cape()
        assert np.array_equal(cape == 0.0, cref == 0.0)
        assert rel(cape, cref) <= 1e-11


def test_coarse_latitude_grid_matches_oracle(oracle):
    """At 37 latitudes the zonal circle factor is taken about the equator row 18 (the
    reference takes row 90, the equator of its 181 rows): two calls against the oracle, with the
    entrainment rate bounded."""
    grid = (21, 37, 73)
    print(f"case moist convection fp64 {grid}", flush=True)
    ref, snap, evap = planted(oracle, grid, 23)
    with A.Integration(grid) as integ:
        integ.load(snap)
        integ.set_conv("Evaporation", evap)
        for call in (1, 2):
            integ.moist_convection()
            ref.moist_convection()
            e = errors(integ, ref)
            bad = {k: v for k, v in e.items() if v > 1e-11}
            assert not bad, (call, bad)
        E = integ.get_conv("E_u")
        assert np.all(np.isfinite(E)) and np.abs(E).max() < 0.05


@pytest.mark.parametrize("order,precision", [("jacobi", "fp64"), ("lex", "fp64"), ("jacobi", "mixed")])
def test_convection_inside_the_step(oracle, order, precision):
    """The configured moist block on the device (athad_set_moist_convection): eight iterations,
    four of them moist, from the planted state with the seeded Evaporation input and seeded MC_*
    tendencies (which the scheme carries where no column convects, into RK4 and the damping),
    against the
    oracle with convection in its moist block; FP64 within 1e-10 relative L2 per field,
    ice-scheme array and convection array (measured values printed), the mixed form within 1e-4
    on the state. Graph replay equals eager launches bit for bit."""
    grid = (21, 37, 73)
    print(f"case convection in step {grid} {order} {precision}", flush=True)
    ref = oracle.OracleState(*grid, seed=31)
    ref.run(2, "jacobi")
    snap = ref.snapshot()
    mp = {k: np.zeros(grid) for k in oracle.MP_FIELDS}
    oracle.plant_microphysics(snap, mp, 32)
    evap = np.random.default_rng(31).uniform(-1.0, 6.0, grid[1:])
    ref = oracle.OracleState(*grid, blank=True)
    ref.load(snap)
    ref.set_iter(2, 2)
    for k, v in mp.items():
        ref.mp(k)[...] = v
    ref.evap()[...] = evap
    mcp = np.zeros((len(ARRAYS),) + grid)
    rng = np.random.default_rng(33)
    for name, lo, hi in (("MC_t", -1e-3, 1e-3), ("MC_q", -1e-6, 1e-6), ("MC_v", -1e-4, 1e-4),
                         ("MC_w", -1e-4, 1e-4), ("P_conv", 0.0, 1e-3)):
        mcp[ARRAYS.index(name)] = rng.uniform(lo, hi, grid)
    for q, name in enumerate(ARRAYS):
        ref.mc(name)[...] = mcp[q]
    ref.set_moist_convection(True)
    ref.run(8, "jacobi" if order == "jacobi" else "lex_planes")
    outs = []
    for graph in (False, True):
        with A.Integration(grid, precision=precision, order=order) as integ:
            integ.load(snap)
            for k, v in mp.items():
                integ.set_mp(k, v)
            integ.set_conv("Evaporation", evap)
            for q, name in enumerate(ARRAYS):
                integ.set_conv(name, mcp[q])
            integ.set_iter(2, 2)
            integ.set_moist_convection(True)
            integ.run(8, graph=graph)
            outs.append({**{n: integ.get(n) for n in A.PROGNOSTIC + ["p_dyn", "p_stat"]},
                         **{"mp/" + n: integ.get_mp(n) for n in ("P_rain", "P_snow", "Precipitation", "S_c")},
                         **{"mc/" + n: integ.get_conv(n) for n in ARRAYS}})
    for n in outs[0]:
        assert np.array_equal(outs[0][n], outs[1][n]), ("graph vs eager", n)
    got = outs[1]
    e = {}
    for n, v in got.items():
        if n.startswith("mp/"):
            r = ref.mp(n[3:])
        elif n.startswith("mc/"):
            r = ref.mc(n[3:])
        else:
            r = ref.field(n)
        e[n] = rel(v, r)
    print({k: f"{v:.1e}" for k, v in e.items() if v > 0})
    if precision == "fp64":
        bad = {k: v for k, v in e.items() if not v <= 1e-10}
    else:
        bad = {k: v for k, v in e.items() if not k.startswith(("mp/", "mc/")) and k not in ("cloud", "ice")
               and not v <= 1e-4}
    assert not bad, bad
    # the tendencies reached RK4 and the damping: they moved t and v against a run without them
    assert np.abs(ref.mc("MC_t")).max() > 0 and np.abs(ref.mc("MC_v")).max() > 0

