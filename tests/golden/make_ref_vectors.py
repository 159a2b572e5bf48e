"""Writes the golden vectors that pin the oracle, from the reference's own compiled code.

Runs only where the reference sources exist (the build container): it drives
oracle/_ref/ref_harness (built by oracle/Makefile.ref from the reference's sources at their
built-in grid 61 x 181 x 361) and stores, per check, the reference's output at a fixed
sample of cells plus per-field sums and L2 norms. The inputs are regenerated by the oracle
in the tests from the same seed, so only outputs are stored.

    python tests/golden/make_ref_vectors.py               # ref_vectors.npz
    python tests/golden/make_ref_vectors.py --radiation   # ref_radiation.npz
    python tests/golden/make_ref_vectors.py --ice         # ref_ice.npz
    python tests/golden/make_ref_vectors.py --conv        # ref_conv.npz
    python tests/golden/make_ref_vectors.py --grids       # ref_grids.npz
    python tests/golden/make_ref_vectors.py --conv-steps  # ref_conv_steps.npz
    python tests/golden/make_ref_vectors.py --turb        # ref_turb.npz

--grids needs the harness built at the benchmark grids by oracle/Makefile.grid (a copy of
the reference's sources with only the grid constants and the spacing changed):
    make -C oracle -f Makefile.grid IM=21 JM=91 KM=181 -j8
    make -C oracle -f Makefile.grid IM=41 JM=181 KM=361 -j8
"""
from __future__ import annotations

import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import oracle as O  # noqa: E402

HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
GRID = (61, 181, 361)
SEED = 1234
SPIN = 2          # Jacobi-order oracle steps that make the operator inputs
NSAMPLE = 512
OPS = ["poisson", "satadj", "damp_moist", "moist_clamp", "densities", "evaporation", "turbulence",
       "damp_turb", "bc", "rk4", "polar", "radial", "project"]
ORDER_FREE = {"poisson", "satadj", "damp_moist", "moist_clamp", "densities", "evaporation",
              "damp_turb", "bc", "polar", "radial", "project"}


def sample_index(n, seed=99):
    rng = np.random.default_rng(seed)
    return np.sort(rng.choice(n, size=NSAMPLE, replace=False))


def run_ref(state_in, mode, wd, n=None, threads=1):
    out = os.path.join(wd, "out.bin")
    cmd = [HARNESS, state_in, out, mode] + ([str(n)] if n is not None else [])
    subprocess.run(cmd, check=True, cwd=wd, stdout=subprocess.DEVNULL,
                   env=dict(os.environ, OMP_NUM_THREADS=str(threads)))
    return O.read_bin(out)


def pack(prefix, d, idx, store):
    for n in O.FIELDS:
        a = d[n].reshape(-1)
        store[f"{prefix}/{n}/sample"] = a[idx]
        store[f"{prefix}/{n}/sum"] = np.array(np.sum(a))
        store[f"{prefix}/{n}/l2"] = np.array(np.sqrt(np.sum(a * a)))


def pack_rad(prefix, rad, idx, idx2, store):
    """Radiation outputs: sampled cells (columns for albedo), sums and L2 norms, the short
    wave per latitude and t_skin whole."""
    for name, ix in (("radiation", idx), ("epsilon", idx), ("albedo", idx2)):
        a = rad[name].reshape(-1)
        store[f"{prefix}/{name}/sample"] = a[ix]
        store[f"{prefix}/{name}/sum"] = np.array(np.sum(a))
        store[f"{prefix}/{name}/l2"] = np.array(np.sqrt(np.sum(a * a)))
    store[f"{prefix}/sw"] = rad["sw"]
    store[f"{prefix}/t_skin"] = np.array(rad["t_skin"])


def radiation_vectors():
    """Mode-2 radiation (MultiLayerRadiation, the 5 % blend of t, updateSkinTemperature): once
    on the Jacobi-spun-up state, and after each of two whole lex iterations from the start."""
    wd = tempfile.mkdtemp(prefix="athad_golden_rad_")
    n = GRID[0] * GRID[1] * GRID[2]
    idx, idx2 = sample_index(n), sample_index(GRID[1] * GRID[2], seed=98)
    out = {"grid": np.array(GRID), "seed": np.array(SEED), "spin": np.array(SPIN), "index": idx,
           "index2d": idx2}
    st = O.OracleState(*GRID, seed=SEED)
    init = os.path.join(wd, "init.bin")
    st.save_bin(init)
    o = os.path.join(wd, "out.bin")
    subprocess.run([HARNESS, init, o, "steps_rad", "2"], check=True, cwd=wd, stdout=subprocess.DEVNULL,
                   env=dict(os.environ, OMP_NUM_THREADS="1"))
    pack("steps2_rad", O.read_bin(o), idx, out)
    pack_rad("steps2_rad", O.read_rad(o + ".rad", GRID), idx, idx2, out)
    print("steps with radiation done", flush=True)
    st.run(SPIN, "jacobi")
    base = os.path.join(wd, "base.bin")
    st.save_bin(base, iter_n=SPIN)
    subprocess.run([HARNESS, base, o, "radiation"], check=True, cwd=wd, stdout=subprocess.DEVNULL,
                   env=dict(os.environ, OMP_NUM_THREADS="8"))
    pack("op/radiation", O.read_bin(o), idx, out)
    pack_rad("op/radiation", O.read_rad(o + ".rad", GRID), idx, idx2, out)
    path = os.path.join(HERE, "ref_radiation.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes")


ICE_STATE = ["t", "c", "cloud", "ice", "gr", "S_c_c"]
ICE_PLANT = 5     # seed of oracle.plant_microphysics for the ice-scheme input


def ice_vectors():
    """The two-category ice scheme (TwoCatIceScheme::run, then the damping of P_rain and
    P_snow) twice in a row on the Jacobi-spun-up state with planted condensate, fluxes and
    temperature spread; the second call starts from the first call's fluxes stored as
    P_rainn / P_snown. Then two whole iterations in the reference's single-thread order with
    the scheme in its place in the moist block, from the same planted state."""
    wd = tempfile.mkdtemp(prefix="athad_golden_ice_")
    n = GRID[0] * GRID[1] * GRID[2]
    idx = sample_index(n)
    out = {"grid": np.array(GRID), "seed": np.array(SEED), "spin": np.array(SPIN), "plant": np.array(ICE_PLANT),
           "index": idx}
    st = O.OracleState(*GRID, seed=SEED)
    st.run(SPIN, "jacobi")
    snap = st.snapshot()
    mp = {k: np.zeros(GRID) for k in O.MP_FIELDS}
    O.plant_microphysics(snap, mp, ICE_PLANT)
    base, o = os.path.join(wd, "base.bin"), os.path.join(wd, "out.bin")
    for call in (1, 2):
        st = O.OracleState(*GRID, blank=True)
        st.load(snap)
        st.set_iter(SPIN, SPIN)
        st.save_bin(base, iter_n=SPIN)
        O.save_mp(base + ".mp", mp)
        subprocess.run([HARNESS, base, o, "ice"], check=True, cwd=wd, stdout=subprocess.DEVNULL,
                       env=dict(os.environ, OMP_NUM_THREADS="8"))
        ref, rmp = O.read_bin(o), O.read_mp(o + ".mp", GRID)
        d = {k: ref[k] for k in ICE_STATE}
        d.update(rmp)
        for name, a in d.items():
            a = a.reshape(-1)
            out[f"call{call}/{name}/sample"] = a[idx]
            out[f"call{call}/{name}/sum"] = np.array(np.sum(a))
            out[f"call{call}/{name}/l2"] = np.array(np.sqrt(np.sum(a * a)))
        # next call: the reference's output state, its fluxes stored as the store keeps them
        snap = {k: ref[k] for k in ref}
        mp = dict(rmp)
        mp["P_rainn"], mp["P_snown"] = rmp["P_rain"].copy(), rmp["P_snow"].copy()
        print(f"ice call {call} done", flush=True)
    # two whole lex iterations with the ice scheme in the moist block (the second is a moist
    # iteration), from the same planted state
    st = O.OracleState(*GRID, seed=SEED)
    st.run(SPIN, "jacobi")
    snap = st.snapshot()
    mp = {k: np.zeros(GRID) for k in O.MP_FIELDS}
    O.plant_microphysics(snap, mp, ICE_PLANT)
    st = O.OracleState(*GRID, blank=True)
    st.load(snap)
    st.set_iter(SPIN, SPIN)
    st.save_bin(base, iter_n=SPIN)
    O.save_mp(base + ".mp", mp)
    subprocess.run([HARNESS, base, o, "steps_ice", "2"], check=True, cwd=wd, stdout=subprocess.DEVNULL,
                   env=dict(os.environ, OMP_NUM_THREADS="1"))
    d = dict(O.read_bin(o))
    d.update(O.read_mp(o + ".mp", GRID))
    for name in O.FIELDS + O.MP_FIELDS:
        a = d[name].reshape(-1)
        out[f"steps2_ice/{name}/sample"] = a[idx]
        out[f"steps2_ice/{name}/sum"] = np.array(np.sum(a))
        out[f"steps2_ice/{name}/l2"] = np.array(np.sqrt(np.sum(a * a)))
    print("steps with the ice scheme done", flush=True)
    path = os.path.join(HERE, "ref_ice.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes")


CONV_EVAP_SEED = 11   # numpy seed of the evaporation diagnostic convection reads


def conv_evaporation(grid):
    """The per-column evaporation input [mm/d] of the moist-convection vectors: uniform in
    [-1, 6), so some columns have none (the surface-flux perturbation then stays zero)."""
    return np.random.default_rng(CONV_EVAP_SEED).uniform(-1.0, 6.0, grid[1:])


def conv_vectors():
    """Moist convection (MoistConvection::run) twice in a row on the Jacobi-spun-up state with
    the ice-scheme vectors' planted condensate (so convective precipitation survives the
    downdraft evaporation), from zeroed convection arrays; the second call reads what the first
    left in the arrays the scheme does not clear (E_u, D_u, the updraft and downdraft winds)."""
    wd = tempfile.mkdtemp(prefix="athad_golden_conv_")
    n = GRID[0] * GRID[1] * GRID[2]
    idx = sample_index(n)
    out = {"grid": np.array(GRID), "seed": np.array(SEED), "spin": np.array(SPIN), "plant": np.array(ICE_PLANT),
           "evap_seed": np.array(CONV_EVAP_SEED), "index": idx}
    st = O.OracleState(*GRID, seed=SEED)
    st.run(SPIN, "jacobi")
    snap = st.snapshot()
    O.plant_microphysics(snap, {k: np.zeros(GRID) for k in O.MP_FIELDS}, ICE_PLANT)
    st = O.OracleState(*GRID, blank=True)
    st.load(snap)
    st.set_iter(SPIN, SPIN)
    base = os.path.join(wd, "base.bin")
    st.save_bin(base, iter_n=SPIN)
    conv_evaporation(GRID).tofile(base + ".evap")
    for call in (1, 2):
        o = os.path.join(wd, f"out{call}.bin")
        subprocess.run([HARNESS, base, o, "conv", str(call)], check=True, cwd=wd, stdout=subprocess.DEVNULL,
                       env=dict(os.environ, OMP_NUM_THREADS="8"))
        ref = np.fromfile(o + ".mc").reshape(len(O.MC_FIELDS), n)
        for q, name in enumerate(O.MC_FIELDS):
            a = ref[q]
            out[f"call{call}/{name}/sample"] = a[idx]
            out[f"call{call}/{name}/sum"] = np.array(np.sum(a))
            out[f"call{call}/{name}/l2"] = np.array(np.sqrt(np.sum(a * a)))
        out[f"call{call}/cape"] = np.fromfile(o + ".cape")   # computeCAPE's im values
        print(f"convection call {call} done", flush=True)
    path = os.path.join(HERE, "ref_conv.npz")
    np.savez_compressed(path, **out)
    print("wrote", path)


CONV_STEPS = 4
CONV_MC_SEED = 33
# convection tendencies and convective precipitation planted at the start of the coupled-step
# vectors (uniform, seeded): the scheme does not clear MC_* where no column convects, so they
# carry through the iterations into RK4 and the moist block's damping (P_conv it clears)
CONV_MC_PLANT = (("MC_t", -1e-3, 1e-3), ("MC_q", -1e-6, 1e-6), ("MC_v", -1e-4, 1e-4), ("MC_w", -1e-4, 1e-4),
                 ("P_conv", 0.0, 1e-3))


def conv_planted_mc(grid):
    """The 32 convection arrays (MC_FIELDS order, [q][i][j][k]) at the start of the coupled
    steps: zero except the planted tendencies and P_conv."""
    out = np.zeros((len(O.MC_FIELDS),) + tuple(grid))
    rng = np.random.default_rng(CONV_MC_SEED)
    for name, lo, hi in CONV_MC_PLANT:
        out[O.MC_FIELDS.index(name)] = rng.uniform(lo, hi, grid)
    return out


def conv_steps_vectors():
    """Four whole lex iterations of the configured moist block, the ice scheme and moist
    convection in their places (cAtmosphereModel.cpp:1230-1263), two of them moist, from the
    Jacobi-spun-up state with the planted condensate and the seeded Evaporation input, ice and
    convection arrays zero at the start. Writes ref_conv_steps.npz: the state, the ice-scheme
    arrays and the convection arrays after the four iterations."""
    wd = tempfile.mkdtemp(prefix="athad_golden_convsteps_")
    n = GRID[0] * GRID[1] * GRID[2]
    idx = sample_index(n)
    out = {"grid": np.array(GRID), "seed": np.array(SEED), "spin": np.array(SPIN), "plant": np.array(ICE_PLANT),
           "evap_seed": np.array(CONV_EVAP_SEED), "steps": np.array(CONV_STEPS), "index": idx}
    st = O.OracleState(*GRID, seed=SEED)
    st.run(SPIN, "jacobi")
    snap = st.snapshot()
    mp = {k: np.zeros(GRID) for k in O.MP_FIELDS}
    O.plant_microphysics(snap, mp, ICE_PLANT)
    st = O.OracleState(*GRID, blank=True)
    st.load(snap)
    st.set_iter(SPIN, SPIN)
    base, o = os.path.join(wd, "base.bin"), os.path.join(wd, "out.bin")
    st.save_bin(base, iter_n=SPIN)
    O.save_mp(base + ".mp", mp)
    conv_evaporation(GRID).tofile(base + ".evap")
    conv_planted_mc(GRID).tofile(base + ".mc")
    subprocess.run([HARNESS, base, o, "steps_conv", str(CONV_STEPS)], check=True, cwd=wd,
                   stdout=subprocess.DEVNULL, env=dict(os.environ, OMP_NUM_THREADS="1"))
    d = dict(O.read_bin(o))
    d.update(O.read_mp(o + ".mp", GRID))
    mc = np.fromfile(o + ".mc").reshape(len(O.MC_FIELDS), n)
    for q, name in enumerate(O.MC_FIELDS):
        d["mc/" + name] = mc[q]
    for name, a in d.items():
        a = np.asarray(a).reshape(-1)
        if a.size != n:      # the 2-D fields (vel_star, t_top_init)
            continue
        out[f"steps/{name}/sample"] = a[idx]
        out[f"steps/{name}/sum"] = np.array(np.sum(a))
        out[f"steps/{name}/l2"] = np.array(np.sqrt(np.sum(a * a)))
    ev = np.fromfile(o + ".evap")
    out["steps/evap"] = ev
    path = os.path.join(HERE, "ref_conv_steps.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes")


# Whole lex-order integrations written by the reference at the benchmark grids and at its
# built-in grid: (grid, seed, steps). 21 x 91 x 181 and 41 x 181 x 361 run on a copy of the
# reference compiled at that grid (oracle/Makefile.grid).
GRID_RUNS = [((21, 91, 181), 77, 10), ((21, 91, 181), 77, 100), ((41, 181, 361), 1234, 100),
             ((61, 181, 361), 1234, 20)]
GRID_FIELDS = O.FIELDS[:11] + ["p_dyn", "p_stat", "nue", "aux_u", "aux_v", "aux_w", "prod", "tke_source",
                               "dis_source", "r_humid"]
GRID_NSAMPLE = 2048


def grid_harness(grid):
    if tuple(grid) == GRID:
        return HARNESS
    return os.path.join(ROOT, "oracle", "_ref", "g%dx%dx%d" % tuple(grid), "ref_harness")


def grid_vectors():
    """Every run of GRID_RUNS in the reference's single-thread order from the nominal initial
    state of its seed (the runs go in parallel, one reference thread each), plus the
    reference's thermodynamic init chain at each grid. Stored per run: 2048 sampled cells of
    every field in GRID_FIELDS, and every field's sum and L2 norm."""
    wd = tempfile.mkdtemp(prefix="athad_golden_grids_")
    out = {}
    procs = []
    for q, (grid, seed, steps) in enumerate(GRID_RUNS):
        h = grid_harness(grid)
        if not os.path.exists(h):
            sys.exit(f"missing {h}: make -C oracle -f Makefile.grid IM={grid[0]} JM={grid[1]} KM={grid[2]}")
        st = O.OracleState(*grid, seed=seed)
        init = os.path.join(wd, f"init{q}.bin")
        st.save_bin(init)
        del st
        o = os.path.join(wd, f"out{q}.bin")
        procs.append((q, o, subprocess.Popen([h, init, o, "steps", str(steps)], cwd=wd, stdout=subprocess.DEVNULL,
                                             env=dict(os.environ, OMP_NUM_THREADS="1"))))
    inits = {}
    for grid in sorted({g for g, _, _ in GRID_RUNS}):
        st = O.OracleState(*grid, thermo_only=True)
        blank = os.path.join(wd, "blank%dx%dx%d.bin" % grid)
        st.save_bin(blank)
        del st
        o = os.path.join(wd, "initref%dx%dx%d.bin" % grid)
        subprocess.run([grid_harness(grid), blank, o, "init"], check=True, cwd=wd, stdout=subprocess.DEVNULL,
                       env=dict(os.environ, OMP_NUM_THREADS="2"))
        ref = O.read_bin(o)
        tag = "%dx%dx%d" % grid
        js = [0, grid[1] // 4, grid[1] // 2, (3 * grid[1]) // 4, grid[1] - 1]
        for name in ["t", "p_stat", "r_humid", "r_dry", "c", "co2"]:
            out[f"{tag}/init/{name}"] = ref[name][:, js, 0]
        out[f"{tag}/init/t_top_init"] = ref["t_top_init"][:, 0]
        out[f"{tag}/init/js"] = np.array(js)
        print("init chain", tag, flush=True)
    for q, o, p in procs:
        if p.wait() != 0:
            sys.exit(f"reference run {GRID_RUNS[q]} failed")
        grid, seed, steps = GRID_RUNS[q]
        ref = O.read_bin(o)
        tag = "%dx%dx%d/seed%d/steps%d" % (grid + (seed, steps))
        n = grid[0] * grid[1] * grid[2]
        idx = np.sort(np.random.default_rng(99).choice(n, size=GRID_NSAMPLE, replace=False))
        out[f"{tag}/index"] = idx
        for name in GRID_FIELDS:
            a = ref[name].reshape(-1)
            out[f"{tag}/{name}/sample"] = a[idx]
            out[f"{tag}/{name}/sum"] = np.array(np.sum(a))
            out[f"{tag}/{name}/l2"] = np.array(np.sqrt(np.sum(a * a)))
        print("run", tag, flush=True)
    out["runs"] = np.array([list(g) + [s, n] for g, s, n in GRID_RUNS])
    path = os.path.join(HERE, "ref_grids.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes")


TURB_MODELS = ["k_epsilon", "k_omega", "k_omega_SST"]
TURB_STEPS = 2


def turb_vectors():
    """The three turbulence closures: for each, the closure selected as
    cAtmosphereModel.cpp:484-489 selects it, TurbulenceAtm::init on the nominal state (its
    boundary-layer profile of k and omega or epsilon, the wall values, compute_sources), then
    two whole iterations in the reference's single-thread order with that closure."""
    wd = tempfile.mkdtemp(prefix="athad_golden_turb_")
    n = GRID[0] * GRID[1] * GRID[2]
    idx = sample_index(n)
    out = {"grid": np.array(GRID), "seed": np.array(SEED), "steps": np.array(TURB_STEPS), "index": idx}
    st = O.OracleState(*GRID, seed=SEED)
    init = os.path.join(wd, "init.bin")
    st.save_bin(init)
    o = os.path.join(wd, "out.bin")
    for model in TURB_MODELS:
        subprocess.run([HARNESS, init, o, "turb:" + model, str(TURB_STEPS)], check=True, cwd=wd,
                       stdout=subprocess.DEVNULL, env=dict(os.environ, OMP_NUM_THREADS="1"))
        pack(f"{model}/init", O.read_bin(o + ".init"), idx, out)
        pack(f"{model}/steps", O.read_bin(o), idx, out)
        print(model, "done", flush=True)
    path = os.path.join(HERE, "ref_turb.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes")


def main():
    if "--conv-steps" in sys.argv[1:]:
        conv_steps_vectors()
        return
    if "--grids" in sys.argv[1:]:
        grid_vectors()
        return
    if not os.path.exists(HARNESS):
        sys.exit("build the reference harness first: make -C oracle -f Makefile.ref")
    if "--conv" in sys.argv[1:]:
        conv_vectors()
        return
    if "--turb" in sys.argv[1:]:
        turb_vectors()
        return
    if "--ice" in sys.argv[1:]:
        ice_vectors()
        return
    if "--radiation" in sys.argv[1:]:
        radiation_vectors()
        return
    if not os.path.exists(HARNESS):
        sys.exit("build the reference harness first: make -C oracle -f Makefile.ref")
    wd = tempfile.mkdtemp(prefix="athad_golden_")
    n = GRID[0] * GRID[1] * GRID[2]
    idx = sample_index(n)
    out = {"grid": np.array(GRID), "seed": np.array(SEED), "spin": np.array(SPIN), "index": idx}

    # 1. thermodynamic init chain: whole columns at five latitudes
    st = O.OracleState(*GRID, thermo_only=True)
    blank = os.path.join(wd, "blank.bin")
    st.save_bin(blank)
    ref = run_ref(blank, "init", wd, threads=8)
    for name in ["t", "p_stat", "r_humid", "r_dry", "c", "co2"]:
        out[f"init/{name}"] = ref[name][:, [0, 45, 90, 135, 180], 0]
    out["init/t_top_init"] = ref["t_top_init"][:, 0]
    del st

    # 2. two whole scoped steps from the initial state, lex order (1 reference thread)
    st = O.OracleState(*GRID, seed=SEED)
    init = os.path.join(wd, "init.bin")
    st.save_bin(init)
    ref = run_ref(init, "steps", wd, n=2, threads=1)
    pack("steps2", ref, idx, out)
    print("steps done", flush=True)

    # 3. one operator at a time on the Jacobi-spun-up state
    st.run(SPIN, "jacobi")
    base = os.path.join(wd, "base.bin")
    st.save_bin(base, iter_n=SPIN)
    for op in OPS:
        ref = run_ref(base, "op:" + op, wd, threads=8 if op in ORDER_FREE else 1)
        pack(f"op/{op}", ref, idx, out)
        print("op", op, flush=True)

    path = os.path.join(HERE, "ref_vectors.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
