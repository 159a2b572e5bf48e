"""Benchmark of the ATHAD_COND time step on one MI355X.

Workload: the FP64 drop-in at the full grid 41 x 181 x 361 (2,678,981 cells) from the
nominal initial state, every scoped operator live (Poisson every 4th iteration, moist block
every 2nd). A step is one scoped iteration of all cells; value = cell-updates/s with the
state resident in HBM. The same line reports the mixed-precision form (with its error
against the FP64 form over the same steps), the lexicographic parity form, the roofline of
the dominant kernel (RK4) from HIP events on the launch stream, and the CPU oracle timed on
this host's cores on a bounded sample of the same workload.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

GRID = (41, 181, 361)
METRIC = ("RK4 cell-updates/s at 41×181×361, mixed vs FP64 at stated error; p99 latency")
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6   # MI355X FP64 vector, published
# Algorithmic bytes per cell of the dominant kernel (k_rk4_rows, FP64): reads 11 current
# fields + 11 stored (*n) fields + p_dyn, nue, p_stat, t_eq = 26 words; writes 11 new
# fields + aux_u/v/w + prod, tke_source, dis_source + the 8 fused stores = 25 words.
RK4_WORDS = 26 + 25
# Whole scoped step, compulsory traffic per cell-update (FP64): 84 words.
STEP_WORDS = 84


def note(msg):
    """A progress line on stderr as each leg of the benchmark ends (stdout holds the JSON)."""
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def run_form(A, precision, order, steps, warmup, seed=1234, profile_steps=0, graph=True):
    """Times `steps` scoped iterations (hipGraph replay of the 4-iteration cycle when `graph`),
    then, if asked, `profile_steps` eager iterations with HIP events around each RK4 launch on
    the library's own stream."""
    n = GRID[0] * GRID[1] * GRID[2]
    with A.Integration(GRID, precision=precision, order=order) as integ:
        integ.init_state(seed)
        integ.run(warmup, graph=graph)
        integ.sync()
        t0 = time.perf_counter()
        integ.run(steps, graph=graph)
        integ.sync()
        dt = time.perf_counter() - t0
        fields = {name: integ.get(name) for name in A.PROGNOSTIC}
        rk4_ms, rk4_n = 0.0, 0
        if profile_steps:
            integ.profile("rk4", True)
            integ.run(profile_steps)
            integ.sync()
            rk4_ms, rk4_n = integ.profile_result("rk4")
            integ.profile("rk4", False)
    return {"ms_per_step": dt / steps * 1e3, "cells_per_s": n * steps / dt, "rk4_ms": rk4_ms,
            "rk4_launches": rk4_n, "fields": fields}


def order_spread(A, steps=100, seed=1234):
    """The three FP64 update orders over the same `steps` iterations from the nominal state:
    lex (the reference's single-thread sweep, timed here), jacobi (the throughput form) and
    symmetric (mirror-exact). Reports, per field, the relative L2 distance of jacobi and
    symmetric from lex: how far the throughput form ends from the reference's order."""
    n = GRID[0] * GRID[1] * GRID[2]
    names = A.PROGNOSTIC + ["p_dyn", "p_stat", "nue"]
    out, fields = {}, {}
    for order in ("lex", "jacobi", "symmetric"):
        with A.Integration(GRID, order=order) as integ:
            integ.init_state(seed)
            integ.sync()
            t0 = time.perf_counter()
            integ.run(steps, graph=True)
            integ.sync()
            dt = time.perf_counter() - t0
            fields[order] = {k: integ.get(k) for k in names}
        out[order] = {"ms_per_step": dt / steps * 1e3, "cells_per_s": n * steps / dt}
    import numpy as np

    def rel(a, b):
        return float(np.sqrt(np.sum((a - b) ** 2)) / max(np.sqrt(np.sum(b * b)), 1e-300))
    for order in ("jacobi", "symmetric"):
        e = {k: rel(fields[order][k], fields["lex"][k]) for k in names}
        out[f"{order}_vs_lex_rel_l2"] = e
        out[f"{order}_vs_lex_max"] = max(e.values())
    out["workload"] = (f"{steps} FP64 iterations at 41x181x361 from seed {seed} in each order; "
                       "the reference's own thread-count spread is ~1e-8 to 6e-6 after 2 iterations "
                       "(README.md:205-208, survey probe)")
    return out


def run_loop(A, precision, steps, warmup, seed=1234, graph=True, convection=False):
    """The iteration with the operators built beside the scoped step inside it: the
    two-category ice scheme in the moist block (with moist convection after it when
    `convection`) and mode-2 radiation with the skin-temperature step after the store
    (run_3D_loop without output)."""
    n = GRID[0] * GRID[1] * GRID[2]
    with A.Integration(GRID, precision=precision) as integ:
        integ.init_state(seed)
        integ.set_ice_scheme(True)
        if convection:
            integ.set_moist_convection(True)
        integ.set_radiation(True)
        integ.run(warmup, graph=graph)
        integ.sync()
        t0 = time.perf_counter()
        integ.run(steps, graph=graph)
        integ.sync()
        dt = time.perf_counter() - t0
        t_skin = integ.radiation_outputs()["t_skin"]
    return {"ms_per_step": dt / steps * 1e3, "cells_per_s": n * steps / dt, "t_skin_end": t_skin}


def run_ice(A, precision, steps, warmup, seed=1234, profile_steps=8, graph=True):
    """The same workload with the two-category ice scheme inside the moist block
    (CategoryIceScheme = 2): its sources in RK4, its fluxes stored every iteration. Times the
    step, then the ice-scheme launches (columns, flux boundaries and damping, clamps) alone."""
    n = GRID[0] * GRID[1] * GRID[2]
    with A.Integration(GRID, precision=precision) as integ:
        integ.init_state(seed)
        integ.set_ice_scheme(True)
        integ.run(warmup, graph=graph)
        integ.sync()
        t0 = time.perf_counter()
        integ.run(steps, graph=graph)
        integ.sync()
        dt = time.perf_counter() - t0
        integ.profile("ice", True)
        integ.run(profile_steps)
        integ.sync()
        ice_ms, ice_n = integ.profile_result("ice")
        integ.profile("ice", False)
    return {"ms_per_step": dt / steps * 1e3, "cells_per_s": n * steps / dt,
            "ice_ms_per_call": ice_ms / max(ice_n, 1), "calls_profiled": ice_n}


def run_closures(A, steps, warmup, seed=1234, profile_steps=8, graph=True):
    """The FP64 step with each turbulence closure (TurbulenceAtm.h:24), each from the closure's
    own TurbulenceAtm::init profile: the iteration, then the turbulence update's launches
    (vel_star and compute_sources, every second iteration) alone."""
    n = GRID[0] * GRID[1] * GRID[2]
    out = {}
    for model in ("k_epsilon", "k_omega", "k_omega_SST"):
        with A.Integration(GRID, precision="fp64") as integ:
            integ.init_state(seed)
            integ.set_turbulence_model(model)
            integ.turbulence_init()
            integ.run(warmup, graph=graph)
            integ.sync()
            t0 = time.perf_counter()
            integ.run(steps, graph=graph)
            integ.sync()
            dt = time.perf_counter() - t0
            integ.profile("turb", True)
            integ.run(profile_steps)
            integ.sync()
            t_ms, t_n = integ.profile_result("turb")
            integ.profile("turb", False)
        out[model] = {"ms_per_step": dt / steps * 1e3, "cells_per_s": n * steps / dt,
                      "turbulence_update_ms": t_ms / max(t_n, 1)}
    return out


def run_next_rows(A, reps=10, seed=1234, warmup=8):
    """Wall time per call of the operators built beside the step, on the FP64 state at this
    grid after `warmup` iterations: mode-2 radiation (with the host skin-temperature step),
    the two-category ice scheme alone, the checkpoint diagnostics and a restart image in
    memory. Each call returns synchronised; radiation's HBM rate counts its 9 words per cell
    (t, c, cloud, ice, co2, p_stat read; t, radiation, epsilon written)."""
    n = GRID[0] * GRID[1] * GRID[2]
    out = {}
    with A.Integration(GRID) as integ:
        integ.init_state(seed)
        integ.run(warmup, graph=True)
        integ.sync()

        def timed(name, fn):
            fn()
            integ.sync()
            t0 = time.perf_counter()
            for _ in range(reps):
                fn()
            integ.sync()
            out[name] = {"ms_per_call": (time.perf_counter() - t0) / reps * 1e3}

        timed("radiation", integ.radiation)
        r = out["radiation"]
        r["hbm_GBps"] = 9 * 8 * n / (r["ms_per_call"] * 1e-3) / 1e9
        r["frac_of_hbm_peak"] = r["hbm_GBps"] / HBM_PEAK_GBS
        timed("ice_scheme", integ.ice_scheme)
        # moist convection: one thread per column; 41 words per cell is its floor (9 fields
        # read, 32 arrays written), the serial column recurrences bound it, not HBM
        timed("moist_convection", integ.moist_convection)
        r = out["moist_convection"]
        r["hbm_GBps_41_words"] = 41 * 8 * n / (r["ms_per_call"] * 1e-3) / 1e9
        timed("residuum", integ.residuum)
        timed("field_minmax", lambda: integ.field_minmax("t"))
        timed("meridional_streamfunction", integ.meridional_streamfunction)
        timed("slice", lambda: integ.get_slice("t", 0, GRID[0] // 2))
        timed("restart_image", integ.save_state)
    out["workload"] = f"FP64 state at 41x181x361 after {warmup} iterations; {reps} calls each"
    return out


def measured_traffic():
    """HBM bytes per k_rk4_rows<double> launch at this grid, as recorded by the FETCH_SIZE and
    WRITE_SIZE counter runs (scaled by the copy-kernel calibration) in the committed profile
    summary; None when that summary is absent."""
    path = os.path.join(ROOT, "profiles", "k_rk4_rows_fp64_41x181x361_hbm_traffic.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def measured_flops():
    """FP64 operations per k_rk4_rows<double> launch at this grid, from the SQ_INSTS_VALU_*_F64
    counter pass recorded in the committed profile summary; None when that summary is absent."""
    path = os.path.join(ROOT, "profiles", "k_rk4_rows_fp64_41x181x361_fp64_ops.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("flop_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def calibrate_copy(mib=512, reps=5):
    """A plain device copy of known size (reads and writes `mib` MiB per launch), the
    reference kernel for scaling the FETCH_SIZE / WRITE_SIZE counters."""
    import torch
    x = torch.ones(mib * (1 << 20) // 8, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    for _ in range(reps):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"calibration": "copy", "bytes_read_per_launch": x.numel() * 8,
                      "bytes_written_per_launch": y.numel() * 8, "launches": reps}))


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline():
    """The reference's CPU path on this host's cores, timed on a bounded sample of the same
    workload (41 x 181 x 361 from the nominal state, FP64). The reference source does not travel
    here; the oracle is its clean-room restatement, pinned to it to ~1e-12 over 100 iterations
    at this grid. Three forms: the faithful one (the reference's single-thread sweep with its
    O(jm + km) distance-to-land scans in the SST closure, RHS_Atm_Turb.cpp:640-671) at one
    thread, as OMP_NUM_THREADS=1 runs it; the same at all cores (its sweep taken plane by plane,
    bit-identical, so the result is still the reference's); and the scan-free restatement (the
    scans' closed form min(i, j, k)) at all cores, a stronger baseline. `value` is the
    faithful form at all cores."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    n = GRID[0] * GRID[1] * GRID[2]
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))

    def timed(threads, scan, order, steps):
        O.set_threads(threads)
        st = O.OracleState(*GRID, seed=1234)
        st.set_sst_scan(scan)
        t0 = time.perf_counter()
        st.run(steps, order)
        el = time.perf_counter() - t0
        O.set_threads(0)
        note(f"cpu {order} threads={threads} scan={scan}: {el / steps:.2f} s/step")
        return {"value": n * steps / el, "unit": "cell-updates/s", "threads": threads, "steps": steps,
                "s_per_step": el / steps}

    one = timed(1, True, "lex", 1)
    full = timed(cores, True, "lex_planes", 2)
    fast = timed(cores, False, "lex_planes", 4)
    return {"value": full["value"], "unit": "cell-updates/s", "cores": cores, "kind": "port",
            "sample": ("the first 2 scoped iterations of 41x181x361 from the nominal state, FP64, the "
                       "reference's order with its SST distance-to-land scans, OpenMP over the cells of "
                       "each hyperplane"),
            "cpu": cpu_model(), "nproc": os.cpu_count(), "omp_num_threads": cores,
            "compiler": "g++ -O2 -std=c++17 -fopenmp -ffp-contract=off (oracle/Makefile)",
            "faithful_1_thread": one, "faithful_all_cores": full, "scan_free_all_cores": fast}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=8)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="skip the mixed and lex forms")
    ap.add_argument("--no-graph", action="store_true", help="time eager launches instead of the hipGraph")
    ap.add_argument("--calibrate-copy", action="store_true", help="only run a plain copy of known size")
    ap.add_argument("--next-rows-only", action="store_true",
                    help="only the per-call times of radiation, the ice scheme, diagnostics and restart")
    ap.add_argument("--loop-only", action="store_true",
                    help="only the FP64 iteration with the ice scheme and radiation inside it")
    ap.add_argument("--ice-only", action="store_true",
                    help="only the FP64 step with the two-category ice scheme in the moist block")
    ap.add_argument("--lex-batched-only", action="store_true",
                    help="only the lex order with 1 and 32 members sharing the plane launches")
    ap.add_argument("--orders-only", action="store_true",
                    help="only the lex / jacobi / symmetric spread over 100 iterations")
    ap.add_argument("--lex-only", action="store_true",
                    help="only the FP64 step in the reference's lex order (the 1e-10 drop-in)")
    ap.add_argument("--closures-only", action="store_true",
                    help="only the FP64 step with each turbulence closure")
    ap.add_argument("--mixed-only", action="store_true",
                    help="only time the mixed form (for profiling it alone); prints its own JSON line")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU sharding is design-only "
                                   "(members or latitude bands would shard across GPUs), --gpus must be 1"}))
        sys.exit(2)
    import numpy as np
    import athad_cond_amd as A

    n = GRID[0] * GRID[1] * GRID[2]
    graph = not args.no_graph
    if args.next_rows_only:
        print(json.dumps({"grid": list(GRID), **run_next_rows(A)}))
        return
    if args.lex_batched_only:
        from athad_cond_amd.ensemble import run_ensemble
        print(json.dumps({f"{m}x{g[0]}x{g[1]}x{g[2]}": run_ensemble(m, g, steps=8, warmup=2, order="lex")
                          for m, g in ((1, (41, 91, 181)), (32, (41, 91, 181)), (32, GRID))}))
        return
    if args.orders_only:
        print(json.dumps({"grid": list(GRID), **order_spread(A, steps=args.steps)}))
        return
    if args.loop_only:
        r = run_loop(A, "fp64", args.steps, args.warmup, graph=graph)
        print(json.dumps({"form": "fp64 with the ice scheme and radiation", "grid": list(GRID),
                          "steps": args.steps, "warmup": args.warmup, **r}))
        return
    if args.ice_only:
        r = run_ice(A, "fp64", args.steps, args.warmup, graph=graph)
        print(json.dumps({"form": "fp64 with the ice scheme", "grid": list(GRID), "steps": args.steps,
                          "warmup": args.warmup, **r}))
        return
    if args.lex_only:
        r = run_form(A, "fp64", "lex", args.steps, args.warmup, graph=graph)
        r.pop("fields")
        print(json.dumps({"form": "fp64 lex", "grid": list(GRID), "steps": args.steps, "warmup": args.warmup, **r}))
        return
    if args.closures_only:
        print(json.dumps({"grid": list(GRID), "steps": args.steps, "warmup": args.warmup,
                          **run_closures(A, args.steps, args.warmup, graph=graph)}))
        return
    if args.mixed_only:
        mixed = run_form(A, "mixed", "jacobi", args.steps, args.warmup, profile_steps=20, graph=graph)
        rk4_s = mixed["rk4_ms"] / max(mixed["rk4_launches"], 1) * 1e-3
        ach = RK4_WORDS * 4 * n / rk4_s / 1e9   # FP32 storage: 4-byte words
        print(json.dumps({"form": "mixed", "grid": list(GRID), "steps": args.steps, "warmup": args.warmup,
                          "ms_per_step": mixed["ms_per_step"], "cells_per_s": mixed["cells_per_s"],
                          "rk4": {"kernel": "k_rk4_jacobi<float>", "avg_launch_ms": rk4_s * 1e3,
                                  "algorithmic_GBps": ach, "frac_of_hbm_peak": ach / HBM_PEAK_GBS}}))
        return
    fp64 = run_form(A, "fp64", "jacobi", args.steps, args.warmup, profile_steps=20, graph=graph)
    note(f"fp64 jacobi: {fp64['ms_per_step']:.3f} ms/step")
    avg_rk4_s = fp64["rk4_ms"] / max(fp64["rk4_launches"], 1) * 1e-3
    achieved = RK4_WORDS * 8 * n / avg_rk4_s / 1e9
    traffic = measured_traffic()
    out = {
        "metric": METRIC,
        "value": fp64["cells_per_s"],
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": fp64["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (nominal initial state, seeded mirrored perturbation)",
        "config": {"workload": "FP64 drop-in, Jacobi order, full scoped step at 41x181x361",
                   "grid": list(GRID), "order": "jacobi", "precision": "fp64", "parallelism": "single GPU",
                   "launch": "hipGraph of 4 iterations" if graph else "eager"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "k_rk4_rows<double>", "avg_launch_ms": avg_rk4_s * 1e3,
                     "bytes_per_cell": RK4_WORDS * 8, "algorithmic_bytes_per_launch": RK4_WORDS * 8 * n},
        "step_hbm_frac": STEP_WORDS * 8 * fp64["cells_per_s"] / (HBM_PEAK_GBS * 1e9),
    }
    flops = measured_flops()
    if flops is not None:
        # the same kernel against the FP64 vector peak: counted FP64 operations per launch over
        # the launch time measured above
        out["alu_roofline"] = {"bound": "fp64 vector", "achieved": flops / avg_rk4_s / 1e12,
                               "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                               "frac": flops / avg_rk4_s / 1e12 / FP64_PEAK_TFLOPS,
                               "kernel": "k_rk4_rows<double>", "flop_per_launch": flops}
    if not args.no_extra:
        mixed = run_form(A, "mixed", "jacobi", args.steps, args.warmup, graph=graph)
        note(f"mixed jacobi: {mixed['ms_per_step']:.3f} ms/step")
        err = {k: float(np.sqrt(np.sum((mixed["fields"][k] - fp64["fields"][k]) ** 2))
                        / max(np.sqrt(np.sum(fp64["fields"][k] ** 2)), 1e-300)) for k in A.PROGNOSTIC}
        out["mixed"] = {"cells_per_s": mixed["cells_per_s"], "ms_per_step": mixed["ms_per_step"],
                        "speedup_vs_fp64": mixed["cells_per_s"] / fp64["cells_per_s"],
                        "rel_l2_vs_fp64": err, "max_rel_l2": max(err.values())}
        from athad_cond_amd.ensemble import run_ensemble
        out["orders"] = order_spread(A, steps=100)
        note("order spread done")
        out["lex"] = {"cells_per_s": out["orders"]["lex"]["cells_per_s"],
                      "ms_per_step": out["orders"]["lex"]["ms_per_step"], "steps": 100}
        # the reference order with members sharing every hyperplane launch
        out["lex_batched"] = {f"{m}x{g[0]}x{g[1]}x{g[2]}": run_ensemble(m, g, steps=8, warmup=2, order="lex")
                              for m, g in ((1, (41, 91, 181)), (32, (41, 91, 181)), (32, GRID))}
        note("batched lex done")
        out["jacobi_vs_lex_rel_l2"] = out["orders"]["jacobi_vs_lex_rel_l2"]
        out["symmetric_vs_lex_rel_l2"] = out["orders"]["symmetric_vs_lex_rel_l2"]
        out["ice_scheme"] = run_ice(A, "fp64", args.steps, args.warmup, graph=graph)
        out["ice_scheme"]["workload"] = ("FP64 Jacobi step at 41x181x361 with the two-category ice scheme "
                                         "in the moist block and its sources in RK4")
        note("ice scheme done")
        out["loop_with_ice_and_radiation"] = run_loop(A, "fp64", args.steps, args.warmup, graph=graph)
        note("loop with ice and radiation done")
        out["loop_with_ice_and_radiation"]["workload"] = (
            "FP64 Jacobi iteration at 41x181x361 with the ice scheme in the moist block and radiation "
            "with the skin-temperature step after the store, all on the device (hipGraph)")
        out["configured_loop"] = run_loop(A, "fp64", max(8, args.steps // 4), 4, graph=graph, convection=True)
        out["configured_loop"]["workload"] = (
            "FP64 Jacobi iteration at 41x181x361 with the ice scheme and moist convection in the moist block "
            "and radiation after the store (run_3D_loop without output), hipGraph")
        note("configured loop done")
        out["turbulence_models"] = run_closures(A, max(8, args.steps // 4), 4, graph=graph)
        out["turbulence_models"]["workload"] = (
            "FP64 Jacobi step at 41x181x361 with each turbulence closure from its TurbulenceAtm::init "
            "profile; turbulence_update_ms: vel_star and compute_sources per call")
        note("turbulence closures done")
        out["next_rows"] = run_next_rows(A)
        note("next rows done")
        out["ensemble"] = {prec: run_ensemble(32, (41, 91, 181), steps=40, precision=prec)
                           for prec in ("fp64", "mixed")}
        note("ensemble done")
        from athad_cond_amd import serving
        reqs = serving.poisson_requests(100, rate_per_s=4.0, seed=99, steps=50)
        server = serving.Server(max_slots=4, chunk=4)
        out["serving"] = server.run(reqs)
        server.close()
        note("serving done")
        out["serving"]["arrivals"] = ("100 requests, Poisson, mean 4 req/s, seed 99; 50 steps each; grids "
                                      "21x91x181, 41x181x361, 61x361x721; mixed with FP64 fallback; p99 is the "
                                      "99th percentile of the 100 latencies (linear interpolation)")
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline()
        note("cpu baseline done")
    print(json.dumps(out))


if __name__ == "__main__":
    main()
