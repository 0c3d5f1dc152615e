#!/usr/bin/env python3
"""Autotuner of the library's launch-shape choices.

For each grid size class (nodes per side 65, 129, 257, 513, 1025) it runs an
argon discharge of that size -- 100 particles per cell per species, argon at
100 mTorr, null-collision MCC, pair re-injection, the mixed-precision
tolerance solve (workloads.with_argon) -- and measures the step time over
the choices the library otherwise takes from csrc/tuned.h:

  sort_key     counting-sort key (tile) width, 5..7 (a new run per value)
  sort interval of each species (passes per counting sort), by coordinate
               descent on the live run (pic_set_tuning)
  mg_sweeps    red-black sweeps per leg of the tolerance solve, 1..3

It writes csrc/tuned.h (the table compiled into libpic2d.so; a species'
best interval k becomes the constant c = k d of the rule k = round(c / d),
d = sigma dt / dx its thermal displacement per step in cells) and tuned.json
(every measurement).  Rebuild the library afterwards (make, or
__graft_entry__.build()).  Results do not depend on any of these choices.

    python autotune.py [--classes 65,129,257,513,1025] [--steps 20]
"""
import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

import pic2d  # noqa: E402
import workloads  # noqa: E402

KEYS = (6, 5, 7)  # the first is the default; another must beat it by ACCEPT
INTERVALS_E = (1, 2, 3, 4, 6, 8)
INTERVALS_I = (4, 8, 16)
SWEEPS = (1, 2, 3)
ACCEPT = 0.01  # a candidate must beat the incumbent by 1 %
D_FLOOR = 0.05  # the library's rule floors the displacement d at 0.05 cells


def log(*a):
    print("[autotune]", *a, file=sys.stderr, flush=True)


def class_params(nodes, ppc):
    return workloads.with_argon(workloads.discharge(nodes - 1, ppc))


def step_ms(sim, steps, reps=2):
    """best of `reps` timings of `steps` fused steps (ms per step)"""
    best = float("inf")
    for _ in range(reps):
        sim.sync()
        t = time.perf_counter()
        sim.advance(steps)
        sim.sync()
        best = min(best, (time.perf_counter() - t) * 1e3 / steps)
    return best


def tune_class(nodes, ppc, steps, warmup):
    rows = []
    best = None
    for key in KEYS:
        p = class_params(nodes, ppc)
        p.sort_key = key
        with pic2d.Pic(p) as sim:
            sim.run(1)
            sim.advance(warmup)
            start = sim.tuning()
            cur = dict(e=start["sort_interval_e"], i=start["sort_interval_i"],
                       nu=start["mg_sweeps"])

            def measure(what, setting):
                sim.set_tuning(setting["e"], setting["i"], setting["nu"])
                ms = step_ms(sim, steps)
                rows.append(dict(nodes=nodes, sort_key=key, sort_interval_e=setting["e"],
                                 sort_interval_i=setting["i"], mg_sweeps=setting["nu"], ms=ms,
                                 sweep=what))
                log(f"N={nodes} key={key} e={setting['e']} i={setting['i']} "
                    f"nu={setting['nu']}: {ms:.4f} ms ({what})")
                return ms

            # coordinate descent; the run's step time drifts as the plasma
            # evolves, so every candidate is timed right after the incumbent
            # and must beat it by more than ACCEPT
            for name, cands in (("e", INTERVALS_E), ("i", INTERVALS_I), ("nu", SWEEPS)):
                for v in cands:
                    if v == cur[name]:
                        continue
                    t_inc = measure("incumbent", cur)
                    cand = dict(cur, **{name: v})
                    t = measure(name, cand)
                    if t < t_inc * (1.0 - ACCEPT):
                        cur = cand
            base = min(measure("final", cur) for _ in range(2))
            if best is None or base < best["ms"] * (1.0 - ACCEPT):
                best = dict(sort_key=key, ms=base, **cur)
            dx = p.length / (p.nodes - 1)
            d = [p.sigma_load_e * p.dt / dx, p.sigma_load_i * p.dti / dx]
    best["sort_c"] = [round(best["e"] * max(d[0], D_FLOOR), 3),
                      round(best["i"] * max(d[1], D_FLOOR), 3)]
    best["displacement"] = d
    log(f"N={nodes}: best {best}")
    return best, rows


HEADER = """// tuned.h -- the library's launch-shape choices per grid size class, as the
// autotuner (autotune.py) measured them; regenerate with it, do not edit.
// A row applies to grids of up to max_nodes nodes per side (the last row to
// every larger grid):
//   sort_c[s]  passes per counting sort of species s (0 electrons, 1 ions)
//              = round(sort_c / d), d = sigma dt / dx the thermal
//              displacement per step in cells, clamped to 1..16
//   sort_key   counting-sort key (tile) width: cell key >> sort_key
//   mg_sweeps  red-black sweeps per leg of the tolerance solve's V-cycle
#pragma once

namespace pic_host {

struct TunedRow {
    int max_nodes;
    double sort_c[2];
    int sort_key;
    int mg_sweeps;
};

// %s
static const TunedRow TUNED[] = {
%s
};

inline const TunedRow &tuned_row(int nodes) {
    const int n = (int)(sizeof(TUNED) / sizeof(TUNED[0]));
    for (int k = 0; k < n; k++)
        if (nodes <= TUNED[k].max_nodes) return TUNED[k];
    return TUNED[n - 1];
}

}  // namespace pic_host
"""


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--classes", default="65,129,257,513,1025")
    ap.add_argument("--ppc", type=int, default=100)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=40)
    ap.add_argument("--out-dir", default=HERE)
    args = ap.parse_args()
    classes = [int(x) for x in args.classes.split(",")]
    results, rows = [], []
    for n in classes:
        best, r = tune_class(n, args.ppc, args.steps, args.warmup)
        results.append(dict(max_nodes=n, **best))
        rows += r
    lines = ",\n".join(
        f"    {{{b['max_nodes']}, {{{b['sort_c'][0]}, {b['sort_c'][1]}}}, {b['sort_key']}, {b['nu']}}}"
        for b in results)
    note = (f"measured by autotune.py on an MI355X (gfx950): argon discharges, {args.ppc} "
            f"particles per cell per species, best of 2 x {args.steps} steps per choice")
    with open(os.path.join(args.out_dir, "csrc", "tuned.h"), "w") as f:
        f.write(HEADER % (note, lines))
    with open(os.path.join(args.out_dir, "tuned.json"), "w") as f:
        json.dump(dict(note=note, best=results, measurements=rows), f, indent=1)
    print(json.dumps(dict(best=results)))


if __name__ == "__main__":
    main()
