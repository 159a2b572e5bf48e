"""The FP64 forms at the full grid 41 x 181 x 361 over 100 iterations (seed 1234).

- lex (the reference's single-thread sweep, RungeKutta_Atm_Turb.cpp:99-264): against the
  oracle's plane-parallel lex sweep, whole fields, relative L2 <= 1e-10 per field; and both
  against the reference's own 100-iteration run at this grid (tests/golden/ref_grids.npz,
  written by a copy of the reference compiled at 41 x 181 x 361).
- jacobi (the throughput form): against the oracle's Jacobi order, <= 1e-10.
- the distance of the jacobi and symmetric forms from the reference order (lex), per field,
  printed; a guard of 1e-3 on the prognostic and pressure fields (measured: t 5e-7, u, v, w
  1e-8 to 6e-8, c 5e-6, tke 6e-5, dis 1e-5, p_dyn 9e-6). Cloud water, ice, nue and aux_u are
  printed only: ice is rounding noise of order 1e-23 whose relative distance is O(1), and
  nue (1.5e-3) follows the threshold switches of the turbulence closure. These forms are
  therefore not 1e-10 drop-ins for the reference's order; lex is.
- mixed precision against the FP64 oracle.
"""
import os

import numpy as np
import pytest

import athad_cond_amd as A
from golden_check import GRIDS, grid_run_errors, rel

pytestmark = pytest.mark.gpu

C2 = (41, 181, 361)
SEED = 1234
STEPS = 100
TOL = 1e-10
PARITY_FIELDS = A.PROGNOSTIC + ["p_dyn", "p_stat", "nue", "aux_u", "aux_v", "aux_w"]
GOLD_FIELDS = PARITY_FIELDS + ["prod", "tke_source", "dis_source", "r_humid"]
# aux_u = rhs_u + dp/dr at the fourth stage (RHS_Atm_Turb.cpp:1077): the sum adds back the
# pressure gradient the tendency subtracted, so it carries the rounding of that gradient
# amplified by |dp/dr| / |aux_u|. The device rounds rhs_u differently from the host (fused
# multiply-adds, device math), so its aux_u differs from the reference by more than the state
# does (measured 2.2e-10 after 100 iterations, with u at 8e-13 and p_dyn at 3e-11; the oracle,
# which keeps the reference's evaluation order, is at 2e-14). Bar for aux_u: 1e-9.
BAR = {"aux_u": 1e-9}


def gpu_run(order, precision="fp64"):
    with A.Integration(C2, precision=precision, order=order) as integ:
        integ.init_state(SEED)
        integ.run(STEPS, graph=True)
        return {n: integ.get(n) for n in GOLD_FIELDS}


@pytest.fixture(scope="module")
def gpu_lex():
    print("case full grid 41x181x361 fp64 lex n=100 (GPU)", flush=True)
    return gpu_run("lex")


@pytest.fixture(scope="module")
def gpu_jacobi():
    print("case full grid 41x181x361 fp64 jacobi n=100 (GPU)", flush=True)
    return gpu_run("jacobi")


@pytest.fixture(scope="module")
def oracle_lex(oracle):
    """The oracle after 100 lex iterations, swept plane by plane on the host threads."""
    print("case oracle 41x181x361 lex n=100", flush=True)
    return run_with_progress(oracle.OracleState(*C2, seed=SEED), "lex_planes", "41x181x361")


@pytest.fixture(scope="module")
def oracle_jacobi(oracle):
    print("case oracle 41x181x361 jacobi n=100", flush=True)
    return run_with_progress(oracle.OracleState(*C2, seed=SEED), "jacobi", "41x181x361")


def run_with_progress(state, order, tag):
    """STEPS iterations of the oracle in chunks of ten, with a line after each chunk."""
    for done in range(0, STEPS, 10):
        state.run(min(10, STEPS - done), order)
        print(f"  oracle {tag} {order}: {min(done + 10, STEPS)}/{STEPS}", flush=True)
    return state


def worst(fields, ref, names):
    return {n: rel(fields[n], ref.field(n)) for n in names}


def test_lex_full_grid_matches_oracle(gpu_lex, oracle_lex):
    w = worst(gpu_lex, oracle_lex, PARITY_FIELDS)
    print("lex vs oracle", {k: f"{v:.2e}" for k, v in w.items()})
    bad = {k: v for k, v in w.items() if not v <= TOL}
    assert not bad, f"fields above {TOL}: {bad}"


@pytest.mark.skipif(not os.path.exists(GRIDS), reason="reference vectors at the benchmark grids not generated")
def test_lex_full_grid_matches_reference(gpu_lex, oracle_lex):
    """The device and the oracle against the reference's own lex run at 41 x 181 x 361
    (2048 sampled cells, sum and L2 of every field)."""
    gold = np.load(GRIDS, allow_pickle=False)
    for tag, get in (("gpu", gpu_lex.__getitem__), ("oracle", oracle_lex.field)):
        e = grid_run_errors(get, gold, C2, SEED, STEPS, GOLD_FIELDS)
        print(tag, "vs reference", {k: f"{v:.2e}" for k, v in e.items()})
        bad = {k: v for k, v in e.items() if not v <= BAR.get(k, TOL)}
        assert not bad, f"{tag}: fields above their bar against the reference: {bad}"


def test_jacobi_full_grid_matches_oracle(gpu_jacobi, oracle_jacobi):
    w = worst(gpu_jacobi, oracle_jacobi, PARITY_FIELDS)
    print("jacobi vs oracle", {k: f"{v:.2e}" for k, v in w.items()})
    bad = {k: v for k, v in w.items() if not v <= TOL}
    assert not bad, f"fields above {TOL}: {bad}"


def test_order_spread_against_reference_order(gpu_lex, gpu_jacobi):
    """How far the throughput form (jacobi) and the mirror-exact form (symmetric) end from the
    reference's order after 100 iterations at the full grid."""
    sym = gpu_run("symmetric")
    for tag, f in (("jacobi", gpu_jacobi), ("symmetric", sym)):
        w = {n: rel(f[n], gpu_lex[n]) for n in PARITY_FIELDS}
        print(f"{tag} vs lex", {k: f"{v:.2e}" for k, v in w.items()})
        guard = ["t", "u", "v", "w", "c", "co2", "tke", "dis", "p_dyn", "p_stat"]
        bad = {k: w[k] for k in guard if not w[k] <= 1e-3}
        assert not bad, f"{tag} vs lex above 1e-3: {bad}"


def test_mixed_precision_error_full_grid(oracle_jacobi):
    """The mixed form after 100 iterations at the full grid against the FP64 oracle: every
    prognostic field, p_dyn, p_stat and nue within 1e-5 relative L2, except cloud water and
    ice, whose errors are bounded by 1e-5 of the total water instead. Cloud water forms as the
    excess of vapour over q_sat(T) in barely supersaturated cells: a difference of two values
    near 0.33 stored in FP32 (relative 6e-8), so a storage rounding of T or c moves cloud by
    about 6e-8 x 0.33 / 8e-5 ~ 2.5e-4 of itself (measured 5.4e-4 relative, 2.2e-7 of the
    total water). Ice is a trace below 3e-13 kg/kg made the same way (measured 1.7e-3 of
    itself, 3e-16 of the total water)."""
    print("case full grid 41x181x361 mixed jacobi n=100", flush=True)
    f = gpu_run("jacobi", precision="mixed")
    fields = ["t", "u", "v", "w", "c", "gr", "co2", "tke", "dis", "p_dyn", "p_stat", "nue"]
    w = worst(f, oracle_jacobi, fields)
    tw = oracle_jacobi.field("c") + oracle_jacobi.field("cloud") + oracle_jacobi.field("ice")
    tw_l2 = float(np.sqrt(np.sum(tw * tw)))
    for n in ("cloud", "ice"):
        d = f[n] - oracle_jacobi.field(n)
        w[n + "_of_total_water"] = float(np.sqrt(np.sum(d * d))) / tw_l2
    extra = worst(f, oracle_jacobi, ["cloud", "ice"])
    print({k: f"{v:.2e}" for k, v in {**w, **extra}.items()})
    bad = {k: v for k, v in w.items() if not v <= 1e-5}
    assert not bad, f"mixed fields above 1e-5: {bad}"
