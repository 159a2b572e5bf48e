"""The oracle's update orders, and the oracle on the benchmark grids against the reference (CPU).

1. The plane-parallel lex sweep (the only way the full grid's lex run fits a test) is bit for
   bit the sequential lex sweep of RungeKutta_Atm_Turb.cpp:99-264 at one thread.
2. The Jacobi restatement reads only the start-of-RK4 neighbourhood: its result does not depend
   on the thread count or on the sweep direction. A read of a value written in the same sweep
   (the lex behaviour) would make both differ; lex and Jacobi differ here, so the checks bite.
3. tests/golden/ref_grids.npz holds the reference's own lex runs at 21 x 91 x 181 (10 and 100
   iterations), 41 x 181 x 361 (100) and its built-in 61 x 181 x 361 (20 iterations), written by copies of the
   reference compiled at each grid (oracle/Makefile.grid), plus its thermodynamic init chain
   at each grid. The oracle must reproduce them to 1e-10 (S_c_c excluded: a diagnostic whose
   Newton break-off flips at rounding level and never feeds back). The 100-iteration runs are
   checked where their oracle runs already happen, beside the device (tests/test_gpu_parity.py
   at 21 x 91 x 181, tests/test_gpu_full_grid.py at 41 x 181 x 361).
"""
import os

import numpy as np
import pytest

from golden_check import GRIDS, grid_run_errors, rel

C1 = (21, 91, 181)
TOL = 1e-10
NAMES = ["t", "u", "v", "w", "c", "cloud", "ice", "gr", "co2", "tke", "dis", "p_dyn", "p_stat", "nue", "aux_u",
         "aux_v", "aux_w", "prod", "tke_source", "dis_source", "r_humid"]


def run(oracle, order, steps=3, threads=0, grid=C1, seed=5):
    oracle.set_threads(threads)
    st = oracle.OracleState(*grid, seed=seed)
    st.run(steps, order)
    snap = st.snapshot()
    oracle.set_threads(0)
    return snap


def test_lex_planes_equal_sequential_lex(oracle):
    a, b = run(oracle, "lex"), run(oracle, "lex_planes")
    for n in a:
        assert np.array_equal(a[n], b[n]), n


def test_jacobi_independent_of_threads_and_direction(oracle):
    one = run(oracle, "jacobi", threads=1)
    many = run(oracle, "jacobi", threads=8)
    rev = run(oracle, "jacobi_reverse", threads=3)
    lex = run(oracle, "lex")
    for n in one:
        assert np.array_equal(one[n], many[n]), ("threads", n)
        assert np.array_equal(one[n], rev[n]), ("direction", n)
    assert not np.array_equal(one["u"], lex["u"])


gold = pytest.mark.skipif(not os.path.exists(GRIDS), reason="reference vectors at the benchmark grids not generated")


@gold
@pytest.mark.parametrize("grid", [(21, 91, 181), (41, 181, 361), (61, 181, 361)])
def test_init_chain_on_grid(oracle, grid):
    g = np.load(GRIDS, allow_pickle=False)
    tag = "%dx%dx%d" % grid
    st = oracle.OracleState(*grid, thermo_only=True)
    js = g[f"{tag}/init/js"]
    for name in ["t", "p_stat", "r_humid", "r_dry", "c", "co2"]:
        assert rel(st.field(name)[:, js, 0], g[f"{tag}/init/{name}"]) < 1e-13, name
    assert rel(st.field2d("t_top_init")[:, 0], g[f"{tag}/init/t_top_init"]) < 1e-13


@gold
@pytest.mark.parametrize("grid,seed,steps", [((21, 91, 181), 77, 10), ((61, 181, 361), 1234, 20)],
                         ids=["21x91x181-10", "61x181x361-20"])
def test_lex_run_on_grid(oracle, grid, seed, steps):
    g = np.load(GRIDS, allow_pickle=False)
    st = oracle.OracleState(*grid, seed=seed)
    st.run(steps, "lex_planes")
    e = grid_run_errors(st.field, g, grid, seed, steps, NAMES)
    print({k: f"{v:.2e}" for k, v in e.items()})
    bad = {k: v for k, v in e.items() if not v <= TOL}
    assert not bad, f"{grid} {steps} steps: {bad}"


@pytest.mark.parametrize("jm", [91, 181])
def test_convection_circle_factor_on_coarse_grids(oracle, jm):
    """Moist convection's zonal circle factor is taken about the equator row of every grid
    (row 90 of the reference's 181 rows, MoistConvection.h:239). Taken about row 90 on a
    91-row grid it would be the pole: the factor would turn negative beyond row 45 and the
    entrainment gradients would blow up where it crosses zero. On the same kind of input the
    entrainment rate stays bounded and of the same size on both grids."""
    grid = (21, jm, 73)
    ref = oracle.OracleState(*grid, seed=21)
    ref.run(2, "jacobi")
    snap = ref.snapshot()
    oracle.plant_microphysics(snap, {k: np.zeros(grid) for k in oracle.MP_FIELDS}, 22)
    ref.load(snap)
    ref.evap()[...] = np.random.default_rng(21).uniform(-1.0, 6.0, grid[1:])
    ref.moist_convection()
    ref.moist_convection()
    E = ref.mc("E_u")
    assert np.all(np.isfinite(E))
    assert 1e-4 < np.abs(E).max() < 0.05, np.abs(E).max()
    assert np.all(np.abs(ref.mc("MC_t")) <= 0.01 + 1e-15)


def test_sst_land_scans_equal_closed_form(oracle):
    """The reference finds the SST wall distance by scanning the land mask along j and k from
    every cell in every stage (RHS_Atm_Turb.cpp:640-671); with no land those scans leave
    min(i, j, k), which the oracle (and the device) use directly. Both forms give the same bits
    over two lex iterations."""
    grid = (11, 21, 41)
    a = oracle.OracleState(*grid, seed=6)
    b = oracle.OracleState(*grid, seed=6)
    b.set_sst_scan(True)
    a.run(2, "lex")
    b.run(2, "lex")
    for n in oracle.FIELDS:
        assert np.array_equal(a.field(n), b.field(n)), n
