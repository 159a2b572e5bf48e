"""GPU radiation (run_3D_loop's radiation_mode 2) against the CPU oracle.

MultiLayerRadiation (MultiLayerRadiation.h:59-384), the omega_rad blend of t
(cAtmosphereModel.cpp:863-879) and updateSkinTemperature (:276-311). The oracle is pinned to
the reference's compiled code by tests/test_oracle_golden.py and by the closed forms of
tests/test_oracle_radiation.py. Tolerances: one call 1e-12 (relative L2; the device forms the
flux sweeps as parallel compositions and T^4 as (T^2)^2, rounding differently from the
sequential loop and pow by a few ulps), FP64 iterations 1e-10, the mixed form per field (its
test says why).
"""
import numpy as np
import pytest

import athad_cond_amd as A

pytestmark = pytest.mark.gpu


def rel(a, b):
    den = np.sqrt(np.sum(b * b))
    num = np.sqrt(np.sum((a - b) ** 2))
    return num / den if den > 0 else num


def check(integ, ref, tol, tag, fields=("t",)):
    out, ro = integ.radiation_outputs(), ref.rad_outputs()
    errs = {n: rel(integ.get(n), ref.field(n)) for n in fields}
    for n in ("radiation", "epsilon", "albedo"):
        assert np.all(np.isfinite(out[n])), f"{tag}: non-finite {n}"
        errs[n] = rel(out[n], ro[n])
    errs["t_skin"] = abs(out["t_skin"] / ro["t_skin"] - 1.0)
    bad = {k: v for k, v in errs.items() if v > tol}
    assert not bad, f"{tag}: above {tol}: {bad}"
    return errs


@pytest.mark.parametrize("grid", [(21, 37, 73), (70, 11, 21)])
def test_one_call_matches_oracle(oracle, grid):
    print(f"case radiation once {grid}", flush=True)
    ref = oracle.OracleState(*grid, seed=7)
    ref.run(4, "jacobi")
    state = ref.snapshot()
    with A.Integration(grid) as integ:
        integ.load(state)
        integ.set_iter(4, ref.total_iter)
        t_before = integ.get("t")
        integ.radiation()
        ref.radiation()
        e = check(integ, ref, 1e-12, f"{grid}")
        print({k: f"{v:.1e}" for k, v in e.items()})
        # the blend moves t by 5 % of the way to radiative equilibrium: a real change
        assert rel(integ.get("t"), t_before) > 1e-6


def test_iterations_with_radiation_match_oracle(oracle):
    grid = (21, 37, 73)
    print(f"case radiation 16 iterations {grid}", flush=True)
    ref = oracle.OracleState(*grid, seed=9)
    with A.Integration(grid) as integ:
        integ.init_state(9)
        for _ in range(16):
            integ.step()
            integ.radiation()
            ref.step("jacobi")
            ref.radiation()
        e = check(integ, ref, 1e-10, "16 iterations", fields=A.PROGNOSTIC + ["p_stat", "r_humid"])
        print({k: f"{v:.1e}" for k, v in e.items()})
        # t_skin has moved away from its configured 254 K and the densities use it
        assert abs(integ.radiation_outputs()["t_skin"] - 254.0) > 1.0


def test_mixed_precision_with_radiation(oracle):
    """The mixed form drifts further from FP64 with radiation than without (t 9e-6 against
    1e-7 after 8 iterations here): the albedo follows the cloud water, the mixed form's least
    accurate field (FP32 saturation adjustment, 3e-3 here), and t_skin, which densities
    impose on the upper levels, follows the planetary albedo. Bars at ten times the measured
    errors; the albedo's is the cloud water's."""
    grid = (21, 37, 73)
    print(f"case radiation mixed {grid}", flush=True)
    ref = oracle.OracleState(*grid, seed=9)
    with A.Integration(grid, precision="mixed") as integ:
        integ.init_state(9)
        for _ in range(8):
            integ.step()
            integ.radiation()
            ref.step("jacobi")
            ref.radiation()
        out, ro = integ.radiation_outputs(), ref.rad_outputs()
        errs = {n: rel(integ.get(n), ref.field(n)) for n in ("t", "c", "p_stat")}
        errs.update({n: rel(out[n], ro[n]) for n in ("radiation", "epsilon", "albedo")})
        errs["t_skin"] = abs(out["t_skin"] / ro["t_skin"] - 1.0)
        print({k: f"{v:.1e}" for k, v in errs.items()})
        bars = {"t": 1e-4, "c": 2e-4, "p_stat": 1e-4, "radiation": 2e-4, "epsilon": 3e-4, "albedo": 2e-2,
                "t_skin": 5e-4}
        bad = {k: v for k, v in errs.items() if v > bars[k]}
        assert not bad, bad


def test_graph_runs_pick_up_the_new_t_skin():
    """Eager steps and captured 4-iteration graphs give the same bits around radiation calls."""
    grid = (11, 21, 41)
    print(f"case radiation graph {grid}", flush=True)
    with A.Integration(grid) as a, A.Integration(grid) as b:
        a.init_state(2)
        b.init_state(2)
        for _ in range(3):
            a.run(4, graph=False)
            b.run(4, graph=True)
            a.radiation()
            b.radiation()
        for n in A.PROGNOSTIC + ["p_stat", "r_humid"]:
            np.testing.assert_array_equal(a.get(n), b.get(n), err_msg=n)
        assert a.radiation_outputs()["t_skin"] == b.radiation_outputs()["t_skin"]


def test_ensemble_members_equal_their_single_runs():
    grids = [(11, 21, 41), (21, 37, 73)]
    print(f"case radiation ensemble {grids}", flush=True)
    with A.Integration(grids) as ens:
        ens.init_state(4)
        ens.run(2)
        ens.radiation()
        for q, g in enumerate(grids):
            with A.Integration(g) as one:
                one.init_state(4 + q)
                one.run(2)
                one.radiation()
                np.testing.assert_array_equal(ens.get("t", member=q), one.get("t"))
                a, b = ens.radiation_outputs(q), one.radiation_outputs()
                for n in ("radiation", "epsilon", "albedo", "t_skin"):
                    np.testing.assert_array_equal(a[n], b[n], err_msg=n)


def test_radiation_errors():
    print("case radiation errors", flush=True)
    with A.Integration((11, 21, 41)) as integ:
        integ.init_state(1)
        with pytest.raises(A.AthadError, match="status 5"):
            integ.radiation_outputs()
    with A.Integration((81, 11, 21)) as integ:
        integ.init_state(1)
        with pytest.raises(A.AthadError, match="status 4"):
            integ.radiation()


@pytest.mark.parametrize("graph", [False, True])
def test_radiation_inside_the_step(oracle, graph):
    """athad_set_radiation: the radiation branch and the skin-temperature step run on the device
    after every iteration's store (hipGraph replay included); the same as calling
    athad_radiation after each step, and as the oracle."""
    grid = (21, 37, 73)
    print(f"case radiation inside the step graph={graph} {grid}", flush=True)
    ref = oracle.OracleState(*grid, seed=9)
    with A.Integration(grid) as integ:
        integ.init_state(9)
        integ.set_radiation(True)
        integ.run(16, graph=graph)
        for _ in range(16):
            ref.step("jacobi")
            ref.radiation()
        e = check(integ, ref, 1e-10, "16 iterations in the step", fields=A.PROGNOSTIC + ["p_stat", "r_humid"])
        print({k: f"{v:.1e}" for k, v in e.items()})
        assert abs(integ.radiation_outputs()["t_skin"] - 254.0) > 1.0
        # switching it off leaves t_skin where the device moved it and stops the updates
        integ.set_radiation(False)
        ts = integ.radiation_outputs()["t_skin"]
        integ.run(4, graph=graph)
        assert integ.radiation_outputs()["t_skin"] == ts


def test_ice_scheme_and_radiation_inside_the_step(oracle):
    """Both next operators inside the step: the moist block with the ice scheme, then after the
    store the radiation; the whole iteration stays on the device."""
    grid = (21, 37, 73)
    print(f"case ice scheme and radiation inside the step {grid}", flush=True)
    ref = oracle.OracleState(*grid, seed=5)
    ref.run(2, "jacobi")
    snap = ref.snapshot()
    mp = {k: np.zeros(grid) for k in oracle.MP_FIELDS}
    oracle.plant_microphysics(snap, mp, 6)
    ref.load(snap)
    for k, v in mp.items():
        ref.mp(k)[...] = v
    ref.set_ice_scheme(True)
    with A.Integration(grid) as integ:
        integ.load(snap)
        integ.set_iter(2, ref.total_iter)
        for k, v in mp.items():
            integ.set_mp(k, v)
        integ.set_ice_scheme(True)
        integ.set_radiation(True)
        integ.run(8, graph=True)
        for _ in range(8):
            ref.step("jacobi")
            ref.radiation()
        e = check(integ, ref, 1e-10, "ice + radiation", fields=A.PROGNOSTIC + ["p_stat", "r_humid"])
        for n in ("P_rain", "P_snow", "S_v"):
            e[n] = rel(integ.get_mp(n), ref.mp(n))
        print({k: f"{v:.1e}" for k, v in e.items()})
        assert max(e[n] for n in ("P_rain", "P_snow", "S_v")) < 1e-10
