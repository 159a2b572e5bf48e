"""GPU two-category ice scheme (CategoryIceScheme = 2) against the CPU oracle.

TwoCatIceScheme::run (TwoCatIceScheme.h:89-111) and run_3D_loop's damping of P_rain and
P_snow (cAtmosphereModel.cpp:1243-1244). The oracle is pinned to the reference's compiled
scheme by tests/test_oracle_golden.py::test_ice_scheme_twice on the same kind of input: a
spun-up state with condensate, falling precipitation and a temperature spread planted by
oracle.plant_microphysics so that every branch of the scheme runs. Tolerance 1e-11 relative
L2 per array in FP64 (the device's pow/exp differ from glibc's by an ulp; measured values are
printed).
"""
import numpy as np
import pytest

import athad_cond_amd as A
from athad_cond_amd import _lib as L

pytestmark = pytest.mark.gpu

STATE = ["t", "c", "cloud", "ice", "gr", "S_c_c"]


def rel(a, b):
    den = np.sqrt(np.sum(b * b))
    num = np.sqrt(np.sum((a - b) ** 2))
    return num / den if den > 0 else num


def planted(oracle, grid, seed, spin=2):
    ref = oracle.OracleState(*grid, seed=seed)
    ref.run(spin, "jacobi")
    snap = ref.snapshot()
    mp = {k: np.zeros(grid) for k in oracle.MP_FIELDS}
    oracle.plant_microphysics(snap, mp, seed + 1)
    ref.load(snap)
    for k, v in mp.items():
        ref.mp(k)[...] = v
    return ref, snap, mp


def load(integ, ref, snap, mp, member=0):
    integ.load(snap, member=member)
    integ.set_iter(2, ref.total_iter, member=member)
    for k, v in mp.items():
        integ.set_mp(k, v, member=member)


def errors(integ, ref, member=0):
    e = {n: rel(integ.get(n, member=member), ref.field(n)) for n in STATE}
    for n in L.MP_FIELDS:
        e[n] = rel(integ.get_mp(n, member=member), ref.mp(n))
    return e


@pytest.mark.parametrize("grid", [(21, 37, 73), (41, 91, 181), (70, 11, 21)])
def test_two_calls_match_oracle(oracle, grid):
    """Grids up to 64 levels run the wave-per-column kernel, taller ones the thread-per-column
    kernel."""
    print(f"case ice scheme fp64 {grid}", flush=True)
    ref, snap, mp = planted(oracle, grid, 11)
    with A.Integration(grid) as integ:
        load(integ, ref, snap, mp)
        for call in (1, 2):
            integ.ice_scheme()
            ref.ice_scheme()
            e = errors(integ, ref)
            print(call, {k: f"{v:.1e}" for k, v in e.items()}, flush=True)
            bad = {k: v for k, v in e.items() if v > 1e-11}
            assert not bad, f"call {call}: {bad}"
            for n in ("P_rain", "P_snow", "S_r", "S_s", "S_i"):
                assert np.abs(ref.mp(n)).max() > 0.0, n
            # the store copies the fluxes for the next call's warm start
            integ.store()
            ref.ice_store()
            np.testing.assert_array_equal(integ.get_mp("P_rainn"), integ.get_mp("P_rain"))


def test_iterations_with_the_ice_scheme(oracle):
    """The scoped iterations with the ice scheme called after each moist iteration's step: the
    step's store keeps P_rainn, P_snown (as storeIntermediateData3D does) for the next call."""
    grid = (21, 37, 73)
    print(f"case ice scheme in 8 iterations {grid}", flush=True)
    ref, snap, mp = planted(oracle, grid, 13)
    with A.Integration(grid) as integ:
        load(integ, ref, snap, mp)
        for it in range(8):
            integ.step()
            ref.step("jacobi")
            if it % 2 == 1:
                integ.ice_scheme()
                ref.ice_scheme()
        e = errors(integ, ref)
        for n in A.PROGNOSTIC:
            e[n] = rel(integ.get(n), ref.field(n))
        print({k: f"{v:.1e}" for k, v in e.items()}, flush=True)
        bad = {k: v for k, v in e.items() if v > 1e-10}
        assert not bad, bad
        # the last call left new fluxes; the next iteration's store copies them
        assert not np.array_equal(integ.get_mp("P_rainn"), integ.get_mp("P_rain"))
        integ.step()
        np.testing.assert_array_equal(integ.get_mp("P_rainn"), integ.get_mp("P_rain"))
        np.testing.assert_array_equal(integ.get_mp("P_snown"), integ.get_mp("P_snow"))


def test_mixed_precision_ice_scheme(oracle):
    """FP32 storage, FP64 arithmetic: the planted inputs round to FP32, so each array moves by
    FP32 rounding plus what that rounding does to the fluxes accumulated down the column.
    Measured: state 4e-8, fluxes and sources 9e-6 (Precipitation); bars at 1e-6 and 1e-4."""
    grid = (21, 37, 73)
    print(f"case ice scheme mixed {grid}", flush=True)
    ref, snap, mp = planted(oracle, grid, 11)
    with A.Integration(grid, precision="mixed") as integ:
        load(integ, ref, snap, mp)
        integ.ice_scheme()
        ref.ice_scheme()
        e = errors(integ, ref)
        print({k: f"{v:.1e}" for k, v in e.items()}, flush=True)
        bad = {k: v for k, v in e.items() if v > (1e-6 if k in STATE else 1e-4)}
        assert not bad, bad


def test_members_equal_their_single_runs(oracle):
    grids = [(21, 37, 73), (11, 21, 41)]
    print(f"case ice scheme ensemble {grids}", flush=True)
    inputs = [planted(oracle, g, 17 + q) for q, g in enumerate(grids)]
    single = []
    for g, (ref, snap, mp) in zip(grids, inputs):
        with A.Integration(g) as integ:
            load(integ, ref, snap, mp)
            integ.ice_scheme()
            single.append({n: integ.get_mp(n) for n in L.MP_FIELDS} | {n: integ.get(n) for n in STATE})
    with A.Integration(grids) as integ:
        for q, (ref, snap, mp) in enumerate(inputs):
            load(integ, ref, snap, mp, member=q)
        integ.ice_scheme()
        for q in range(len(grids)):
            for n in L.MP_FIELDS:
                np.testing.assert_array_equal(integ.get_mp(n, member=q), single[q][n], err_msg=f"{q} {n}")
            for n in STATE:
                np.testing.assert_array_equal(integ.get(n, member=q), single[q][n], err_msg=f"{q} {n}")


def test_arrays_read_zero_before_first_use():
    with A.Integration((11, 21, 41)) as integ:
        integ.init_state(3)
        for n in L.MP_FIELDS:
            assert not integ.get_mp(n).any()
        buf = np.empty((11, 21, 41))
        ptr = buf.ctypes.data_as(L.ctypes.POINTER(L.ctypes.c_double))
        assert L.lib().athad_get_mp_field(integ.h, 0, L.MP_FIELDS.index("P_rain") + 99, ptr) == 1   # ATHAD_ERR_ARG
        assert L.lib().athad_get_mp_field(integ.h, 5, 0, ptr) == 1


@pytest.mark.parametrize("order,grid,steps", [("jacobi", (21, 37, 73), 8), ("lex", (11, 21, 41), 4)])
def test_ice_scheme_inside_the_step(oracle, order, grid, steps):
    """CategoryIceScheme = 2 inside the moist block: SatAdj with S_c_c uncapped, damping of t
    ice c cloud at every level, the scheme and its flux damping, precipitationSum, cap_S, the
    clamps; the sources in RK4, their boundaries, falling precipitation evaporated by SatAdj.
    Steps through the hipGraph too (athad_run), which must have been recaptured for it."""
    print(f"case ice scheme in the step {order} {grid} {steps}", flush=True)
    ref, snap, mp = planted(oracle, grid, 19)
    ref.set_ice_scheme(True)
    with A.Integration(grid, order=order) as integ:
        load(integ, ref, snap, mp)
        integ.run(4, graph=True)     # graph of the scoped step captured, then dropped below
        ref.set_ice_scheme(False)
        ref.run(4, order)
        integ.set_ice_scheme(True)
        ref.set_ice_scheme(True)
        integ.run(steps, graph=True)
        ref.run(steps, order)
        e = errors(integ, ref)
        for n in A.PROGNOSTIC + ["p_stat", "r_humid", "p_dyn"]:
            e[n] = rel(integ.get(n), ref.field(n))
        print({k: f"{v:.1e}" for k, v in e.items()}, flush=True)
        bad = {k: v for k, v in e.items() if v > (1e-8 if k == "S_c_c" else 1e-10)}
        assert not bad, bad
        for n in ("P_rain", "P_snow", "S_v", "S_r", "S_s"):
            assert np.abs(ref.mp(n)).max() > 0.0, n


def test_standalone_call_between_graph_runs(oracle):
    """With the scheme in the step the fluxes are stored right after it; a standalone call in
    between changes them outside the step, so the next iterations store them in full (and the
    captured graph is rebuilt). Checked against the oracle doing the same sequence."""
    grid = (21, 37, 73)
    print(f"case ice scheme standalone between graph runs {grid}", flush=True)
    ref, snap, mp = planted(oracle, grid, 23)
    ref.set_ice_scheme(True)
    with A.Integration(grid) as integ:
        load(integ, ref, snap, mp)
        integ.set_ice_scheme(True)
        integ.run(8, graph=True)
        ref.run(8, "jacobi")
        integ.ice_scheme()
        ref.ice_scheme()
        integ.run(9, graph=True)
        ref.run(9, "jacobi")
        e = errors(integ, ref)
        for n in A.PROGNOSTIC:
            e[n] = rel(integ.get(n), ref.field(n))
        print({k: f"{v:.1e}" for k, v in e.items()}, flush=True)
        bad = {k: v for k, v in e.items() if v > (1e-8 if k == "S_c_c" else 1e-10)}
        assert not bad, bad
        np.testing.assert_array_equal(integ.get_mp("P_rainn"), integ.get_mp("P_rain"))
        np.testing.assert_array_equal(integ.get_mp("P_snown"), integ.get_mp("P_snow"))


def test_symmetric_form_stays_mirror_exact_with_the_ice_scheme(oracle):
    """The ice scheme is column-local and its boundaries, damping, clamps and RK4 sources are
    pointwise or use the symmetric form's paired latitude sums, so a mirrored state stays
    mirrored bit for bit (v, vn, aux_v change sign) through iterations with the scheme inside
    the moist block."""
    grid = (21, 37, 73)
    print(f"case ice scheme symmetric {grid}", flush=True)
    anti = {"v", "vn", "aux_v"}
    with A.Integration(grid, order="symmetric") as integ:
        integ.init_state(31)
        integ.run(2)
        snap = {n: integ.get(n) for n in A.FIELDS}
        mp = {k: np.zeros(grid) for k in L.MP_FIELDS}
        oracle.plant_microphysics(snap, mp, 37)
        jm = grid[1]
        for d in (snap, mp):
            for n, a in d.items():
                if n in anti:
                    continue
                a[:, jm - 1 - np.arange(jm // 2), :] = a[:, np.arange(jm // 2), :]   # south := north, mirrored
        integ.load(snap)
        for k, v in mp.items():
            integ.set_mp(k, v)
        integ.set_ice_scheme(True)
        integ.run(16, graph=True)
        bad = {}
        for n in A.FIELDS + L.MP_FIELDS:
            x = integ.get_mp(n) if n in L.MP_FIELDS else integ.get(n)
            m = x[:, ::-1, :]
            ok = np.array_equal(x, -m) if n in anti else np.array_equal(x.view(np.uint64), m.view(np.uint64))
            if not ok:
                bad[n] = float(np.abs(x - (-m if n in anti else m)).max())
        assert not bad, bad
        assert np.abs(integ.get_mp("P_rain")).max() > 0.0


def test_members_at_different_phases_with_the_ice_scheme(oracle):
    """Members whose iteration counters differ run the moist block in different launches' turns:
    the cadence gates of the scheme's kernels decide per member, and the fluxes are stored every
    iteration. Each member still equals its own single run bit for bit."""
    grids = [(21, 37, 73), (11, 21, 41)]
    print(f"case ice scheme members at different phases {grids}", flush=True)
    inputs = [planted(oracle, g, 41 + q) for q, g in enumerate(grids)]
    phases = [2, 3]
    single = []
    for q, (g, (ref, snap, mp)) in enumerate(zip(grids, inputs)):
        with A.Integration(g) as integ:
            load(integ, ref, snap, mp)
            integ.set_iter(phases[q], phases[q])
            integ.set_ice_scheme(True)
            integ.run(6, graph=True)
            single.append({n: integ.get(n) for n in A.PROGNOSTIC} | {n: integ.get_mp(n) for n in L.MP_FIELDS})
    with A.Integration(grids) as integ:
        for q, (ref, snap, mp) in enumerate(inputs):
            load(integ, ref, snap, mp, member=q)
            integ.set_iter(phases[q], phases[q], member=q)
        integ.set_ice_scheme(True)
        integ.run(6, graph=True)
        for q in range(len(grids)):
            for n in A.PROGNOSTIC:
                np.testing.assert_array_equal(integ.get(n, member=q), single[q][n], err_msg=f"{q} {n}")
            for n in L.MP_FIELDS:
                np.testing.assert_array_equal(integ.get_mp(n, member=q), single[q][n], err_msg=f"{q} {n}")


def test_full_grid_loop_with_ice_scheme_and_radiation(oracle):
    """41x181x361 from the nominal initial state, 20 iterations with the ice scheme in the moist
    block and radiation after each store, through the hipGraph, against the oracle doing the
    same: every prognostic field, the fluxes and sources, and t_skin within 1e-10."""
    grid = (41, 181, 361)
    print(f"case full-grid loop with ice scheme and radiation {grid}", flush=True)
    ref = oracle.OracleState(*grid, seed=1234)
    ref.set_ice_scheme(True)
    with A.Integration(grid) as integ:
        integ.init_state(1234)
        integ.set_ice_scheme(True)
        integ.set_radiation(True)
        integ.run(20, graph=True)
        for _ in range(20):
            ref.step("jacobi")
            ref.radiation()
        e = {n: rel(integ.get(n), ref.field(n)) for n in A.PROGNOSTIC + ["p_stat", "r_humid", "p_dyn"]}
        for n in L.MP_FIELDS:
            e[n] = rel(integ.get_mp(n), ref.mp(n))
        e["t_skin"] = abs(integ.radiation_outputs()["t_skin"] / ref.scalar("t_skin") - 1.0)
        print({k: f"{v:.1e}" for k, v in e.items()}, flush=True)
        bad = {k: v for k, v in e.items() if v > 1e-10}
        assert not bad, bad
