#!/usr/bin/env python3
"""Throughput of the fused 2-D PIC timestep on one MI355X.

Prints ONE JSON line on stdout:
  value          particle-steps/s = (N_e + N_i) * steps / wall time, inputs
                 resident in HBM, timed between two stream synchronisations
  roofline       HBM roofline of the dominant kernel (the fused push + scatter +
                 deposit pass, k_push_scatter): algorithmic bytes per launch
                 (48 B per particle + 16 B per grid node) / its mean duration,
                 measured with HIP events on the library's own stream
  cpu_baseline   the CPU oracle in the reference's layout (FP64 48-byte AoS,
                 the reference's phase order) on a bounded sample of the same
                 workload, timed on this host
  latency_ms     p50 / p99 of per-step durations (HIP events around each step)
  accuracy       phi / E error of the split-precision solver against the
                 library's FP64 multigrid solve of the same charge density
                 (converged to 1e-12), reported every 100 steps of an untimed
                 continuation of the run

Workloads "ensemble" (64 independent runs, the C4 mix) and "serving" (Poisson
arrivals at five offered loads, continuous batching under a 200 GB budget)
report the batched path against the plain one-run-at-a-time path and the
per-step latency percentiles instead of a kernel roofline.

Progress goes to stderr.  `--gpus N` with N > 1 exits: this build runs on one GPU.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "vectorized-particle-in-cell-code-with-graphics_amd")
sys.path.insert(0, PKG)

import numpy as np  # noqa: E402

import pic2d  # noqa: E402
import workloads  # noqa: E402

HBM_PEAK_GBS = 8000.0
CPU_SAMPLE = 1 << 24  # particles per species of the CPU baseline at most


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def make_params(name):
    table = {
        "argon_256": lambda: workloads.argon_256(),
        "plain": lambda: workloads.plain(),
        "plain_256": lambda: workloads.discharge(256, 200),
        "plain_1024": lambda: workloads.large_1024(),
        "argon_1024": lambda: workloads.argon_1024(),
    }
    if name not in table:
        raise SystemExit(f"unknown workload {name}; choose from {sorted(table)}")
    return table[name]()


def describe(name, p, tuning):
    nu = tuning["mg_sweeps"]
    return {
        "workload": name,
        "nodes": int(p.nodes),
        "cells": int(p.nodes - 1) ** 2,
        "particles_per_species": int(p.init_particle_no),
        "length_m": p.length,
        "collisions": bool(p.have_collisions),
        "walls": "absorbing" if p.wall == pic2d.WALL_ABSORB else "steady-state plop",
        "mode": "physical" if p.mode == pic2d.MODE_PHYSICAL else "reference",
        "multigrid": ("fp64 FMG ncycle=%d" % p.mg_ncycle) if p.mg_mode == pic2d.MG_FIXED
        else ("fp32 red-black V(%d,%d) cycles + fp64 refinement, tol %g"
              % (nu, nu, p.mg_tol)),
        # launch shapes in effect (explicit, or the library's tuned table)
        "tuning": tuning,
    }


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(p, e, i, steps, threads):
    """Time the oracle's FP64 reference-layout step loop on the host (with the
    FP64 physical-mode collisions and pair re-injection when the workload has
    them)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import orc  # the oracle: only the CPU baseline leg uses it

    dx = p.dx
    N = p.nodes
    _, cref = orc.configs_from_params(p)
    ea = orc.state_to_aos(e, dx)
    ia = orc.state_to_aos(i, dx)
    Ex = np.zeros((N + 1) * (N + 1))
    used = orc.lib().orc_set_threads(threads)
    if p.have_collisions:
        me, mi = orc.mcc_params(p, True), orc.mcc_params(p, False)
        plop = p.wall == pic2d.WALL_STEADY
        nid = [len(ea), len(ia)]
        t0 = time.perf_counter()
        r = orc.run_ref_mcc(cref, me, mi, plop, p.vte, p.vti, steps, ea, ia, e["id"], i["id"],
                            nid, Ex, Ex.copy(), grow=1.5)
        dt = time.perf_counter() - t0
        nbar = 0.5 * (len(ea) + len(ia) + len(r["e"]) + len(r["i"]))
        return nbar * steps / dt, dt, used
    t0 = time.perf_counter()
    orc.run_ref(cref, steps, ea, ia, e["id"], i["id"], Ex, Ex.copy())
    dt = time.perf_counter() - t0
    return (len(ea) + len(ia)) * steps / dt, dt, used


def calibration_copy(nbytes):
    """Three device-to-device copies of nbytes through the HIP runtime's blit
    kernel: a known number of bytes read and written for scaling HBM byte
    counters."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    for ptr in (a, b):
        if hip.hipMalloc(C.byref(ptr), C.c_size_t(nbytes)) != 0:
            raise RuntimeError("hipMalloc failed")
    hip.hipMemset(a, 1, C.c_size_t(nbytes))
    for _ in range(3):
        hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3)  # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    log(f"calibration copy: 3 x {nbytes >> 20} MiB read + {nbytes >> 20} MiB written")


def copy_ceiling(nbytes=1 << 30, reps=5):
    """Device-to-device copy bandwidth through the HIP runtime's blit kernel
    (read + written bytes / time, HIP events): the measured ceiling beside
    the nominal 8 TB/s peak."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    for ptr in (a, b):
        if hip.hipMalloc(C.byref(ptr), C.c_size_t(nbytes)) != 0:
            raise RuntimeError("hipMalloc failed")
    ev0, ev1 = C.c_void_p(), C.c_void_p()
    hip.hipEventCreate(C.byref(ev0))
    hip.hipEventCreate(C.byref(ev1))
    hip.hipMemset(a, 1, C.c_size_t(nbytes))
    hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3)  # warm: hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    hip.hipEventRecord(ev0, None)
    for _ in range(reps):
        hip.hipMemcpyAsync(b, a, C.c_size_t(nbytes), 3, None)
    hip.hipEventRecord(ev1, None)
    hip.hipEventSynchronize(ev1)
    ms = C.c_float()
    hip.hipEventElapsedTime(C.byref(ms), ev0, ev1)
    hip.hipEventDestroy(ev0)
    hip.hipEventDestroy(ev1)
    hip.hipFree(a)
    hip.hipFree(b)
    return 2.0 * nbytes * reps / (ms.value * 1e-3) / 1e9


def live_count(sims):
    return sum(s.particle_count(0) + s.particle_count(1) for s in sims)


def run_ensemble(args):
    """C4: the same 64 runs advanced one at a time (plain) and all at once
    (batched: one host thread and HIP stream per run)."""
    runs = workloads.ensemble(args.runs)
    log(f"ensemble: {len(runs)} runs, creating")
    sims = [pic2d.Pic(p) for p in runs]
    for s in sims:
        s.run(0)  # pre-loop wall pass
    log("warmup", args.warmup, "steps (batched)")
    pic2d.batch_advance(sims, args.warmup)
    n0 = live_count(sims)
    log("plain:", args.steps, "steps per run, one run at a time")
    t0 = time.perf_counter()
    for s in sims:
        s.advance(args.steps)
        s.sync()
    t_plain = time.perf_counter() - t0
    n1 = live_count(sims)
    log("batched:", args.steps, "steps per run, all runs at once")
    t0 = time.perf_counter()
    pic2d.batch_advance(sims, args.steps)
    t_batch = time.perf_counter() - t0
    n2 = live_count(sims)
    plain = 0.5 * (n0 + n1) * args.steps / t_plain
    batched = 0.5 * (n1 + n2) * args.steps / t_batch
    cells = sorted({int(p.nodes - 1) for p in runs})
    for s in sims:
        s.close()
    return {
        "metric": "particle-steps/s",
        "value": batched,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * t_batch / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (Philox-seeded uniform positions, Maxwellian velocities)",
        "config": {"workload": "ensemble", "runs": len(runs), "cells_per_side": cells,
                   "particles": int(n1), "hw_queues": os.environ.get("GPU_MAX_HW_QUEUES")},
        "plain": {"value": plain, "ms_per_step": 1e3 * t_plain / args.steps},
        "batched_vs_plain": batched / plain,
        "roofline": None,
        "cpu_baseline": None,
    }


def run_serving(args):
    """C5: capacity under saturation, then the same requests arriving as a
    Poisson process at 0.25-1.25x that capacity (unit-rate arrival times
    scaled by 1/rate, so every load replays identical work)."""
    budget = int(200e9)
    base = workloads.serving_requests(1.0, 1e9, seed=11, step_scale=args.step_scale,
                                      max_requests=args.requests)
    unit = [t for _, _, t in base]  # arrival times of a unit-rate Poisson process
    sat = [(p, n, 0.0) for p, n, _ in base]
    log(f"serving: capacity run, {len(sat)} requests at t = 0")
    _, st = pic2d.serve(sat, budget, args.max_active)
    cap = st["throughput"]
    per_req = st["particle_steps"] / len(sat)
    loads = []
    for k, load in enumerate((0.25, 0.5, 0.75, 1.0, 1.25)):
        rate = load * cap / per_req
        reqs = [(p, n, t / rate) for (p, n, _), t in zip(base, unit)]
        log(f"serving: load {load}, {len(reqs)} requests at {rate:.2f}/s")
        res, st = pic2d.serve(reqs, budget, args.max_active)
        wait = [r["admit_s"] - q[2] for r, q in zip(res, reqs)]
        phi_err = max(r["phi_error"] for r in res)
        e_err = max(r["e_error"] for r in res)
        loads.append({"offered_load": load, "requests": len(reqs), "rate_per_s": rate,
                      "throughput": st["throughput"], "step_p50_ms": st["step_p50_ms"],
                      "step_p99_ms": st["step_p99_ms"], "queue_wait_mean_s": float(np.mean(wait)),
                      "max_concurrent": st["max_concurrent"], "peak_bytes": st["peak_bytes"],
                      "max_request_phi_rel_l2": phi_err, "max_request_e_rel_l2": e_err})
    full = [l for l in loads if l["offered_load"] == 1.0][0]
    return {
        "metric": "particle-steps/s",
        "value": full["throughput"],
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": None,
        "warmup": 0,
        "ms_per_step": full["step_p50_ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (Philox-seeded uniform positions, Maxwellian velocities)",
        "config": {"workload": "serving", "hbm_budget_bytes": budget,
                   "requests_per_load": args.requests, "step_scale": args.step_scale,
                   "hw_queues": os.environ.get("GPU_MAX_HW_QUEUES")},
        "capacity": cap,
        "latency_ms": {"p50": full["step_p50_ms"], "p99": full["step_p99_ms"]},
        "loads": loads,
        "roofline": None,
        "cpu_baseline": None,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="argon_1024")
    ap.add_argument("--accuracy-steps", type=int, default=100,
                    help="untimed steps after the profile with the phi/E accuracy report "
                         "every 100 steps (0: none)")
    ap.add_argument("--profile-steps", type=int, default=50)
    ap.add_argument("--cpu-steps", type=int, default=0, help="0: choose ~10-30 s of CPU work")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--calibration-copy", action="store_true",
                    help="end with three 256 MiB device copies (a reference point for HBM "
                         "byte counters)")
    ap.add_argument("--mg", choices=["default", "fixed", "tol"], default="default",
                    help="override the workload's multigrid mode")
    ap.add_argument("--sort-interval", type=int, default=0,
                    help="particle passes per counting sort (0: the library default)")
    ap.add_argument("--sort-key", type=int, default=-1,
                    help="counting-sort granularity: cell key >> k (0 cells, 6 8x8 tiles; "
                         "-1: the library default)")
    ap.add_argument("--mg-sweeps", type=int, default=-1,
                    help="tolerance mode: red-black sweeps per V-cycle leg, 1-3 (-1: the "
                         "library default)")
    ap.add_argument("--null-bound", choices=["binned", "global"], default="binned",
                    help="electron null-collision bound: per energy bin (default) or global")
    ap.add_argument("--graphs", action="store_true",
                    help="replay the solve's V-cycle batches from captured hipGraphs instead "
                         "of launching them kernel by kernel")
    ap.add_argument("--runs", type=int, default=64, help="ensemble: number of runs")
    ap.add_argument("--requests", type=int, default=24, help="serving: requests per load")
    ap.add_argument("--step-scale", type=float, default=0.2,
                    help="serving: request lengths are U[500, 5000] steps times this")
    ap.add_argument("--max-active", type=int, default=64, help="serving: most runs at once")
    ap.add_argument("--hw-queues", type=int, default=0,
                    help="ensemble/serving: hardware queues for the runs' streams (1-32; "
                         "0: the runtime's default, 4, measured best for the ensemble)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(runs are independent; see DESIGN.md for how the timestep would shard)",
              file=sys.stderr)
        sys.exit(2)
    if args.workload in ("ensemble", "serving"):
        # more hardware queues let more runs' streams overlap; set before the
        # HIP runtime starts (the library's first call)
        if args.hw_queues > 0:
            os.environ["GPU_MAX_HW_QUEUES"] = str(min(32, args.hw_queues))
        out = run_ensemble(args) if args.workload == "ensemble" else run_serving(args)
        print(json.dumps(out), flush=True)
        return

    p = make_params(args.workload)
    if args.mg != "default":
        p.mg_mode = pic2d.MG_FIXED if args.mg == "fixed" else pic2d.MG_TOL
    if args.sort_interval > 0:
        p.sort_interval = args.sort_interval
    if args.sort_key >= 0:
        p.sort_key = args.sort_key
    if args.mg_sweeps > 0:
        p.mg_sweeps = args.mg_sweeps
    p.null_bound = 1 if args.null_bound == "global" else 0
    p.graphs = 1 if args.graphs else 0
    log("workload", args.workload, "nodes", p.nodes, "particles/species", p.init_particle_no)
    sim = pic2d.Pic(p)
    e0 = i0 = None
    stride = 1
    if not args.no_cpu_baseline:
        # the CPU leg runs on every stride-th particle of the loaded state (at
        # most CPU_SAMPLE per species) on the full grid: its FP64 AoS copies
        # of a 1024^2 run would otherwise take tens of GB of host memory
        stride = max(1, -(-p.init_particle_no // CPU_SAMPLE))
        e0 = {k: v[::stride].copy() for k, v in sim.get_state(0).items()}
        i0 = {k: v[::stride].copy() for k, v in sim.get_state(1).items()}
    tuning = sim.tuning()
    sim.run(0)  # pre-loop wall pass of pic_run
    # the first step from the freshly loaded (unsorted) state, timed alone
    sim.sync()
    t0 = time.perf_counter()
    sim.advance(1)
    sim.sync()
    cold_ms = 1e3 * (time.perf_counter() - t0)
    log("warmup", args.warmup, "steps")
    sim.advance(args.warmup)
    sim.sync()
    n0 = sim.particle_size(0) + sim.particle_size(1)
    log("timing", args.steps, "steps")
    t0 = time.perf_counter()
    sim.advance(args.steps)
    sim.sync()
    t1 = time.perf_counter()
    n1 = sim.particle_size(0) + sim.particle_size(1)
    wall = t1 - t0
    nbar = 0.5 * (n0 + n1)
    value = nbar * args.steps / wall
    ms_per_step = 1e3 * wall / args.steps

    log("profiling", args.profile_steps, "steps with HIP events")
    ne_a, ni_a = sim.particle_size(0), sim.particle_size(1)
    mg0 = sim.mg_stats()
    sim.profile(True)
    sim.advance(args.profile_steps)
    prof = sim.profile_read()
    steps_ms = sim.profile_steps()
    sim.profile(False)
    ne_b, ni_b = sim.particle_size(0), sim.particle_size(1)
    mg1 = sim.mg_stats()
    npart_per_launch = 0.25 * (ne_a + ne_b + ni_a + ni_b)
    alg_bytes = 48.0 * npart_per_launch + 16.0 * p.nodes * p.nodes
    # per-species particle passes: sort passes (k_push_scatter) and in-place
    # passes (k_push_inplace); the dominant one by total time is the roofline
    kern = {}
    for name, tot, cnt in (("k_push_scatter", prof[2], prof[7]), ("k_push_inplace", prof[6], prof[8])):
        if cnt > 0:
            ms = tot / cnt
            kern[name] = {"launches": int(cnt), "mean_launch_ms": ms, "total_ms": tot,
                          "achieved_GBs": round(alg_bytes / (ms * 1e-3) / 1e9, 1)}
    dominant = max(kern, key=lambda k: kern[k]["total_ms"])
    k_ms = kern[dominant]["mean_launch_ms"]
    achieved = alg_bytes / (k_ms * 1e-3) / 1e9
    traffic = None
    tf = os.path.join(ROOT, "profiles", f"traffic_{args.workload}.json")
    if os.path.exists(tf):
        try:
            traffic = json.load(open(tf)).get("kernels", {}).get(dominant, {}).get("bytes_per_launch")
        except Exception:
            traffic = None
    roof = {
        "kernel": dominant,
        "bound": "hbm",
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": traffic,
        # the push + gather + deposit phase as a whole (both species, sort
        # passes, collisions and commits included) against the same peak
        "phase_frac": None,
        "alg_bytes_per_launch": alg_bytes,
        "mean_launch_ms": k_ms,
    }
    steps_ms = np.asarray(steps_ms)
    accuracy = None
    if args.accuracy_steps > 0:
        every = min(100, args.accuracy_steps)
        log(f"accuracy: {args.accuracy_steps} untimed steps, phi/E against the FP64 multigrid "
            f"every {every}")
        sim.set_accuracy_report(every, 1e-12)
        sim.advance(args.accuracy_steps)
        sim.sync()
        rows = sim.accuracy_report()
        sim.set_accuracy_report(0)
        accuracy = {
            "reference": "library FP64 multigrid (full multigrid + FP64 V(1,1) cycles) on the "
                         "same right-hand side, relative residual <= 1e-12",
            "every_steps": every,
            "rows": [{"step": int(r[0]), "solver_relres": r[1], "phi_rel_l2": r[2],
                      "ex_rel_l2": r[3], "ey_rel_l2": r[4], "reference_relres": r[5]}
                     for r in rows],
            "max_phi_rel_l2": float(rows[:, 2].max()) if len(rows) else None,
            "max_e_rel_l2": float(rows[:, 3:5].max()) if len(rows) else None,
        }
    k3n = max(prof[7], 1.0)
    push_ms = prof[1] / args.profile_steps
    phase = {
        "push_phase_ms_per_step": push_ms,
        "kernels": kern,
        "count_kernel_ms": prof[4] / k3n,
        "scan_ms": prof[5] / k3n,
        "sort_interval": [tuning["sort_interval_e"], tuning["sort_interval_i"]],
        "field_solve_ms_per_step": prof[3] / args.profile_steps,
        "mg_cycles_per_step": (mg1[1] - mg0[1]) / args.profile_steps
        if p.mg_mode == pic2d.MG_TOL else p.mg_ncycle,
        "mg_relres_last": mg1[2],
        # push + gather + deposit roofline: 48 B/particle + 16 B/node per
        # species pass over the whole push phase (sort, collisions and
        # commits included)
        "push_phase_frac_of_hbm_peak": round(
            (48.0 * 2 * npart_per_launch + 16.0 * 2 * p.nodes ** 2) /
            (push_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
    }
    # multigrid roofline: about 50 B per node per level visit of a V-cycle
    # (smoothing, residual, restriction, prolongation; 4/3 N^2 nodes over the
    # levels) and 40 B per finest node per FP64 refinement pass; the fixed
    # FMG visits (4/3)^2 N^2 nodes ncycle times.  The field phase's time
    # (rho, right-hand side, solve, E) is the denominator: latency-bound.
    nn = float(p.nodes) ** 2
    if p.mg_mode == pic2d.MG_TOL:
        mg_bytes = phase["mg_cycles_per_step"] * (50.0 * nn * 4.0 / 3.0 + 40.0 * nn)
    else:
        mg_bytes = 50.0 * p.mg_ncycle * (4.0 / 3.0) ** 2 * nn
    field_ms = phase["field_solve_ms_per_step"]
    phase["mg_alg_bytes_per_step"] = mg_bytes
    phase["mg_frac_of_hbm_peak"] = (round(mg_bytes / (field_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
                                    if field_ms > 0 else None)
    phase["cold_first_step_ms"] = cold_ms
    out = {
        "metric": "particle-steps/s",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (Philox-seeded uniform positions, Maxwellian velocities)",
        "config": describe(args.workload, p, tuning),
        "latency_ms": {"p50": float(np.percentile(steps_ms, 50)) if steps_ms.size else None,
                       "p99": float(np.percentile(steps_ms, 99)) if steps_ms.size else None},
        "roofline": roof,
        "phases": phase,
        "accuracy": accuracy,
        "cpu_baseline": None,
    }
    roof["phase_frac"] = phase["push_phase_frac_of_hbm_peak"]
    try:
        ceil = copy_ceiling()
        roof["copy_ceiling_GBs"] = round(ceil, 1)
        roof["frac_of_copy_ceiling"] = round(achieved / ceil, 4)
    except Exception as e:  # the ceiling is a reference figure, not the metric
        log("copy ceiling not measured:", e)
        roof["copy_ceiling_GBs"] = None
    sim.close()

    if not args.no_cpu_baseline:
        n_all = len(e0["id"]) + len(i0["id"])
        threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
        # size the sample from one timed step: about 12 s per thread setting
        _, t_probe, _ = cpu_baseline(p, e0, i0, 1, 1)
        steps = args.cpu_steps or max(1, min(50, int(12.0 / max(t_probe, 1e-3))))
        log(f"cpu baseline: {steps} steps of the FP64 reference-layout oracle, 1 and "
            f"{threads} threads")
        v1, t_1, _ = cpu_baseline(p, e0, i0, steps, 1)
        vn, t_n, used = cpu_baseline(p, e0, i0, steps, threads)
        what = "FP64 AoS reference layout" + (
            " with FP64 MCC and pair re-injection" if p.have_collisions else "")
        out["cpu_baseline"] = {
            "value": vn,
            "unit": "particle-steps/s",
            "cores": used,
            "kind": "port",
            "sample": (f"{steps} steps of the full workload ({n_all} particles)" if stride == 1
                       else f"{steps} steps of every {stride}-th particle of the workload's "
                            f"initial state ({n_all} particles) on its full grid") +
                      f", {what}, OpenMP {used} threads {t_n:.1f} s; 1 thread {t_1:.1f} s",
            "single_thread": {"value": v1, "cores": 1},
            "host_cpu": cpu_model(),
        }
    print(json.dumps(out), flush=True)
    if args.calibration_copy:
        calibration_copy(256 << 20)


if __name__ == "__main__":
    main()
