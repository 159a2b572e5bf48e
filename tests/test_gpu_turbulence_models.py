"""GPU turbulence closures (TurbulenceAtm.h:24: k-epsilon, k-omega, k-omega SST) against the oracle.

The oracle's three closures are pinned to the reference's own runs of each
(tests/test_oracle_golden.py::test_turbulence_closures: TurbulenceAtm::init, then two whole
iterations). Here the device, with the closure chosen by athad_set_turbulence_model, runs
TurbulenceAtm::init (athad_turbulence_init) and then whole iterations, compared with the oracle
doing the same on the same seeded state. Bars: 1e-12 relative L2 after the init, 1e-10 after the
iterations (the FP64 drop-in's bar; the device's exp / pow differ from glibc's by an ulp).
"""
import numpy as np
import pytest

import athad_cond_amd as A
from athad_cond_amd import _lib as L

pytestmark = pytest.mark.gpu

GRID = (21, 37, 73)
MODELS = ["k_epsilon", "k_omega", "k_omega_SST"]
TURB = ["tke", "dis", "nue", "prod", "tke_source", "dis_source", "tken", "disn"]
STATE = A.PROGNOSTIC + ["p_dyn", "p_stat", "nue", "prod", "tke_source", "dis_source"]


def rel(a, b):
    den = np.sqrt(np.sum(b * b))
    num = np.sqrt(np.sum((a - b) ** 2))
    return num / den if den > 0 else num


def errors(integ, ref, names):
    out = {}
    for n in names:
        g = integ.get(n)
        assert np.all(np.isfinite(g)), n
        out[n] = rel(g, ref.field(n))
    return out


@pytest.mark.parametrize("order", ["lex", "jacobi"])
@pytest.mark.parametrize("model", MODELS)
def test_closure_init_and_steps_match_oracle(oracle, model, order):
    print(f"case turbulence {model} {order} {GRID}", flush=True)
    ref = oracle.OracleState(*GRID, seed=7)
    ref.set_turb_model(model)
    ref.turbulence_init()
    with A.Integration(GRID, order=order) as integ:
        integ.init_state(7)
        integ.set_turbulence_model(model)
        integ.turbulence_init()
        e = errors(integ, ref, TURB)
        print("init", {k: f"{v:.1e}" for k, v in e.items()})
        bad = {k: v for k, v in e.items() if v > 1e-12}
        assert not bad, bad
        ref.run(8, order)
        integ.run(8)
        e = errors(integ, ref, STATE)
        print("steps", {k: f"{v:.1e}" for k, v in e.items()})
        bad = {k: v for k, v in e.items() if v > 1e-10}
        assert not bad, bad


def test_closures_give_different_fields():
    """The closure reaches the kernels: after the same init state and iterations, the k-epsilon,
    k-omega and SST fields differ (dis is epsilon in the first, omega in the others)."""
    out = {}
    for model in MODELS:
        with A.Integration(GRID) as integ:
            integ.init_state(7)
            integ.set_turbulence_model(model)
            integ.turbulence_init()
            integ.run(4)
            out[model] = {n: integ.get(n) for n in ("dis", "dis_source", "nue")}
    assert np.linalg.norm(out["k_epsilon"]["dis"]) < 1e-3 * np.linalg.norm(out["k_omega"]["dis"])
    assert rel(out["k_omega"]["dis_source"], out["k_omega_SST"]["dis_source"]) > 1e-3
    assert rel(out["k_omega"]["nue"], out["k_omega_SST"]["nue"]) > 1e-6


def test_switch_after_graph_capture(oracle):
    """A graph captured under one closure is dropped when the closure changes: the next graph
    replays run the new closure's kernels (checked against the oracle), bit-equal to eager steps."""
    print(f"case turbulence switch {GRID}", flush=True)
    ref = oracle.OracleState(*GRID, seed=9)
    ref.run(4, "jacobi")
    ref.set_turb_model("k_omega")
    ref.run(8, "jacobi")
    with A.Integration(GRID) as g, A.Integration(GRID) as e:
        for integ, graph in ((g, True), (e, False)):
            integ.init_state(9)
            integ.run(4, graph=graph)          # SST; the graph run captures a graph
            integ.set_turbulence_model("k_omega")
            integ.run(8, graph=graph)
        for n in STATE:
            assert np.array_equal(g.get(n), e.get(n)), n
        err = errors(g, ref, STATE)
        print({k: f"{v:.1e}" for k, v in err.items()})
        bad = {k: v for k, v in err.items() if v > 1e-10}
        assert not bad, bad


@pytest.mark.parametrize("model", ["k_epsilon", "k_omega"])
def test_mixed_precision_closure(oracle, model):
    """The mixed form with each added closure stays near the FP64 oracle after 8 iterations
    (FP32 storage; bar 1e-4 relative L2 on the prognostic fields and nue)."""
    print(f"case turbulence mixed {model}", flush=True)
    ref = oracle.OracleState(*GRID, seed=7)
    ref.set_turb_model(model)
    ref.turbulence_init()
    ref.run(8, "jacobi")
    with A.Integration(GRID, precision="mixed") as integ:
        integ.init_state(7)
        integ.set_turbulence_model(model)
        integ.turbulence_init()
        integ.run(8)
        names = ["t", "u", "v", "w", "c", "co2", "tke", "dis", "nue"]
        e = errors(integ, ref, names)
        print({k: f"{v:.1e}" for k, v in e.items()})
        bad = {k: v for k, v in e.items() if v > 1e-4}
        assert not bad, bad


@pytest.mark.parametrize("model", ["k_epsilon", "k_omega"])
def test_closure_in_the_configured_moist_block(oracle, model):
    """A closure other than SST with the ice scheme and moist convection live in the moist block:
    from a spun-up state with planted condensate, a seeded Evaporation input and seeded MC_*
    tendencies, the closure's own init, then eight Jacobi iterations, four of them moist, against
    the oracle configured the same way (state, ice-scheme fluxes and convection arrays within
    1e-10 relative L2)."""
    print(f"case turbulence {model} configured loop {GRID}", flush=True)
    ref = oracle.OracleState(*GRID, seed=31)
    ref.run(2, "jacobi")
    snap = ref.snapshot()
    mp = {k: np.zeros(GRID) for k in oracle.MP_FIELDS}
    oracle.plant_microphysics(snap, mp, 32)
    evap = np.random.default_rng(31).uniform(-1.0, 6.0, GRID[1:])
    conv = L.CONV_FIELDS[:-1]
    rng = np.random.default_rng(33)
    mcp = {n: np.zeros(GRID) for n in conv}
    for name, lo, hi in (("MC_t", -1e-3, 1e-3), ("MC_q", -1e-6, 1e-6), ("MC_v", -1e-4, 1e-4),
                         ("MC_w", -1e-4, 1e-4), ("P_conv", 0.0, 1e-3)):
        mcp[name] = rng.uniform(lo, hi, GRID)
    ref = oracle.OracleState(*GRID, blank=True)
    ref.load(snap)
    for k, v in mp.items():
        ref.mp(k)[...] = v
    ref.evap()[...] = evap
    for n, v in mcp.items():
        ref.mc(n)[...] = v
    ref.set_turb_model(model)
    ref.turbulence_init()
    ref.set_iter(2, 2)
    ref.set_moist_convection(True)
    ref.run(8, "jacobi")
    with A.Integration(GRID) as integ:
        integ.load(snap)
        for k, v in mp.items():
            integ.set_mp(k, v)
        integ.set_conv("Evaporation", evap)
        for n, v in mcp.items():
            integ.set_conv(n, v)
        integ.set_turbulence_model(model)
        integ.turbulence_init()
        integ.set_iter(2, 2)
        integ.set_moist_convection(True)
        integ.run(8, graph=True)
        e = errors(integ, ref, STATE)
        e.update({"mp/" + n: rel(integ.get_mp(n), ref.mp(n)) for n in ("P_rain", "P_snow", "S_c", "S_v")})
        e.update({"mc/" + n: rel(integ.get_conv(n), ref.mc(n)) for n in conv})
        print({k: f"{v:.1e}" for k, v in e.items()})
        bad = {k: v for k, v in e.items() if not v <= 1e-10}
        assert not bad, bad
        # convection was live: the scheme rewrote the tendencies it was given
        assert rel(ref.mc("MC_t"), mcp["MC_t"]) > 1e-3 and np.abs(ref.mc("MC_t")).max() > 0


def test_unknown_closure_is_refused():
    with A.Integration(GRID) as integ:
        with pytest.raises(Exception):
            integ.set_turbulence_model(3)
        with pytest.raises(ValueError):
            integ.set_turbulence_model("laminar")
