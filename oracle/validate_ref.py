"""Pins the CPU oracle against the reference's own code (oracle/_ref/ref_harness).

TEST INFRASTRUCTURE ONLY. Runs in the build container, where /root/reference exists; the
reference harness is compiled at the reference's built-in grid (61 x 181 x 361), so every
comparison runs there. For each operator the oracle and the reference start from the same
state (written by the oracle) and the per-field relative L2 difference is reported; then
whole scoped steps are compared (lex order = the reference at OMP_NUM_THREADS=1).

    python oracle/validate_ref.py [--steps N] [--out report.json]
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import oracle as O  # noqa: E402

HARNESS = os.path.join(HERE, "_ref", "ref_harness")
GRID = (61, 181, 361)
ORDER_FREE = {"poisson", "satadj", "damp_moist", "moist_clamp", "densities", "evaporation",
              "damp_turb", "bc", "polar", "radial", "store"}


def run_ref(state_in, mode, n=None, threads=1, workdir=None):
    out = os.path.join(workdir, "ref_out.bin")
    cmd = [HARNESS, state_in, out, mode] + ([str(n)] if n is not None else [])
    env = dict(os.environ, OMP_NUM_THREADS=str(threads))
    subprocess.run(cmd, check=True, env=env, cwd=workdir, stdout=subprocess.DEVNULL)
    return O.read_bin(out)


def compare(ora: dict, ref: dict, names=None):
    res = {}
    for n in names or O.FIELDS + O.FIELDS2D:
        a, b = ora[n], ref[n]
        res[n] = {"rel_l2": O.rel_l2(a, b), "max_abs": float(np.max(np.abs(a - b)))}
    return res


def worst(res, skip=("S_c_c",)):
    return max(v["rel_l2"] for k, v in res.items() if k not in skip)


ICE_STATE = ["t", "c", "cloud", "ice", "gr", "S_c_c"]


def check_ice(st, iter_n, base, wd, plant=None):
    """TwoCatIceScheme twice in a row (the second from the first call's fluxes, stored as
    P_rainn / P_snown), the oracle against the reference on the same inputs; per-field relative
    L2. plant: seed of oracle.plant_microphysics applied to the input first, else zero fluxes."""
    shape = st.shape
    out = {}
    mp_in = None
    snap = st.snapshot()
    if plant is not None:
        mp_in = {n: np.zeros(shape) for n in O.MP_FIELDS}
        O.plant_microphysics(snap, mp_in, plant)
    for call in (1, 2):
        t0 = time.time()
        s2 = O.OracleState(*shape, blank=True)
        s2.load(snap)
        s2.set_iter(iter_n, st.total_iter)
        if mp_in is not None:
            for n, v in mp_in.items():
                s2.mp(n)[...] = v
        s2.save_bin(base, iter_n=iter_n)
        O.save_mp(base + ".mp", s2.mp_snapshot())
        s2.ice_scheme()
        ref = run_ref(base, "ice", threads=8, workdir=wd)
        rmp = O.read_mp(os.path.join(wd, "ref_out.bin.mp"), shape)
        res = compare(s2.snapshot(), ref, ICE_STATE)
        omp = s2.mp_snapshot()
        for n in O.MP_FIELDS:
            res[n] = {"rel_l2": O.rel_l2(omp[n], rmp[n]), "max_abs": float(np.max(np.abs(omp[n] - rmp[n]))),
                      "norm": float(np.sqrt(np.sum(rmp[n] ** 2)))}
        out[f"call{call}"] = res
        print(f"ice scheme call {call}: worst rel L2 {max(v['rel_l2'] for v in res.values()):.3e}  "
              f"|P_rain| {res['P_rain']['norm']:.3e} |P_snow| {res['P_snow']['norm']:.3e}  ({time.time()-t0:.1f}s)",
              flush=True)
        s2.ice_store()
        mp_in = s2.mp_snapshot()
        snap = s2.snapshot()
        del s2
    return out


def check_ice_steps(n, wd, plant=5, spin=2):
    """n lex iterations with the ice scheme inside the moist block (and its sources in the
    RHS), from the spun-up state with planted microphysics; the reference at one thread."""
    t0 = time.time()
    st = O.OracleState(*GRID, seed=1234)
    st.run(spin, "jacobi")
    snap = st.snapshot()
    mp = {k: np.zeros(GRID) for k in O.MP_FIELDS}
    O.plant_microphysics(snap, mp, plant)
    s2 = O.OracleState(*GRID, blank=True)
    s2.load(snap)
    s2.set_iter(spin, spin)
    for k, v in mp.items():
        s2.mp(k)[...] = v
    s2.set_ice_scheme(True)
    base = os.path.join(wd, "base.bin")
    s2.save_bin(base, iter_n=spin)
    O.save_mp(base + ".mp", mp)
    s2.run(n, "lex")
    ref = run_ref(base, "steps_ice", n=n, threads=1, workdir=wd)
    rmp = O.read_mp(os.path.join(wd, "ref_out.bin.mp"), GRID)
    res = compare(s2.snapshot(), ref)
    omp = s2.mp_snapshot()
    for name in O.MP_FIELDS:
        res[name] = {"rel_l2": O.rel_l2(omp[name], rmp[name]), "max_abs": float(np.max(np.abs(omp[name] - rmp[name])))}
    print(f"{n} steps with the ice scheme: worst rel L2 {worst(res):.3e}  ({time.time()-t0:.1f}s)", flush=True)
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=4)
    ap.add_argument("--out", default=os.path.join(HERE, "_ref", "validation.json"))
    ap.add_argument("--spin", type=int, default=6, help="oracle steps before the per-operator checks")
    ap.add_argument("--ice-only", action="store_true", help="only the ice-scheme checks")
    args = ap.parse_args()
    if not os.path.exists(HARNESS):
        sys.exit("reference harness not built: make -C oracle -f Makefile.ref")
    report = {"grid": GRID}
    wd = tempfile.mkdtemp(prefix="athad_val_")
    im, jm, km = GRID
    if args.ice_only:
        st = O.OracleState(im, jm, km, seed=1234)
        st.run(args.spin, "lex")
        report["ice"] = check_ice(st, args.spin, os.path.join(wd, "base.bin"), wd)
        report["ice_planted"] = check_ice(st, args.spin, os.path.join(wd, "base.bin"), wd, plant=5)
        report["ice_steps"] = check_ice_steps(args.steps, wd)
        for k, v in report["ice_steps"].items():
            print(f"   {k:14s} {v['rel_l2']:.3e}  max {v['max_abs']:.3e}")
        return

    # 1. thermodynamic init chain
    t0 = time.time()
    st = O.OracleState(im, jm, km, thermo_only=True)
    ora = st.snapshot()
    blank = os.path.join(wd, "blank.bin")
    st.save_bin(blank)
    ref = run_ref(blank, "init", threads=8, workdir=wd)
    names = ["t", "p_stat", "r_humid", "r_dry", "c", "co2", "t_top_init"]
    report["init"] = compare(ora, ref, names)
    print(f"init chain: worst rel L2 {worst(report['init']):.3e}  ({time.time()-t0:.1f}s)", flush=True)
    del st

    # 2. per-operator checks on a spun-up state
    st = O.OracleState(im, jm, km, seed=1234)
    st.run(args.spin, "lex")
    base = os.path.join(wd, "base.bin")
    iter_n = args.spin
    report["ops"] = {}
    for op in ["poisson", "satadj", "damp_moist", "moist_clamp", "densities", "evaporation",
               "turbulence", "damp_turb", "bc", "rk4", "polar", "radial", "store"]:
        t0 = time.time()
        s2 = O.OracleState(im, jm, km, blank=True)
        s2.load(st.snapshot())
        s2.set_iter(iter_n, st.total_iter)
        s2.save_bin(base, iter_n=iter_n)
        s2.op(op, "lex")
        ref = run_ref(base, "op:" + op, threads=8 if op in ORDER_FREE else 1, workdir=wd)
        report["ops"][op] = compare(s2.snapshot(), ref)
        print(f"op {op:12s}: worst rel L2 {worst(report['ops'][op]):.3e}  ({time.time()-t0:.1f}s)", flush=True)
        del s2

    # 2b. mode-2 radiation once on the same state (MLR, the 5 % blend of t, t_skin)
    t0 = time.time()
    s2 = O.OracleState(im, jm, km, blank=True)
    s2.load(st.snapshot())
    s2.set_iter(iter_n, st.total_iter)
    s2.save_bin(base, iter_n=iter_n)
    s2.radiation()
    ref = run_ref(base, "radiation", threads=8, workdir=wd)
    rr, ro = O.read_rad(os.path.join(wd, "ref_out.bin.rad"), (im, jm, km)), s2.rad_outputs()
    report["radiation"] = compare(s2.snapshot(), ref, ["t"])
    for name in ("radiation", "epsilon", "albedo", "sw"):
        report["radiation"][name] = {"rel_l2": O.rel_l2(ro[name], rr[name]),
                                     "max_abs": float(np.max(np.abs(ro[name] - rr[name])))}
    report["radiation"]["t_skin"] = {"rel_l2": abs(ro["t_skin"] / rr["t_skin"] - 1.0), "max_abs": abs(ro["t_skin"] - rr["t_skin"])}
    print(f"radiation       : worst rel L2 {max(v['rel_l2'] for v in report['radiation'].values()):.3e}  "
          f"({time.time()-t0:.1f}s)", flush=True)
    del s2

    # 2c. the two-category ice scheme (+ the damping of its fluxes), twice: from zero stored
    # fluxes, then from the first call's fluxes stored as P_rainn / P_snown
    report["ice"] = check_ice(st, iter_n, base, wd)
    report["ice_planted"] = check_ice(st, iter_n, base, wd, plant=5)
    report["ice_steps"] = check_ice_steps(args.steps, wd)
    del st

    # 3. whole scoped steps from the initial state (lex = 1 reference thread)
    t0 = time.time()
    st = O.OracleState(im, jm, km, seed=1234)
    init = os.path.join(wd, "init.bin")
    st.save_bin(init)
    ref = run_ref(init, "steps", n=args.steps, threads=1, workdir=wd)
    st.run(args.steps, "lex")
    report["steps"] = {"n": args.steps, "fields": compare(st.snapshot(), ref)}
    print(f"{args.steps} steps: worst rel L2 {worst(report['steps']['fields']):.3e}  ({time.time()-t0:.1f}s)",
          flush=True)
    for k, v in report["steps"]["fields"].items():
        print(f"   {k:12s} {v['rel_l2']:.3e}  max {v['max_abs']:.3e}")
    with open(args.out, "w") as f:
        json.dump(report, f, indent=1)
    print("report:", args.out)


if __name__ == "__main__":
    main()
