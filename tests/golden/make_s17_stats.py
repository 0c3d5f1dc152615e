"""Writes s17_stats.json: the statistics of the reference's only statistical
fixture, its batch-mode example output ("Example Output for Batch Mode/":
REF-mode steady-state argon, 257x257 nodes, 262 144 particles per species,
62 405 steps, graph averages over steps 52 000 - 62 399, timeData every 100
steps).  Data only; run where the reference tree is present:

    python tests/golden/make_s17_stats.py /root/reference

Recorded: the averaged centre-row and centre-column rho, phi and E
(rho_phi_EgridX.txt / rho_phi_EgridY.txt, written by graphs.cpp:1167-1186),
the centre rows of the averaged densities (density2d_e.txt, density2d_i1.txt)
and the timeData probe statistics over the rows of steps >= 52 000 (columns
t ne ni phi(P1) phi(P2) phi(P3) q_e(P3) q_i(P3), pic.cpp:1167-1180).
"""
import json
import os
import sys

import numpy as np

src = os.path.join(sys.argv[1] if len(sys.argv) > 1 else "/root/reference",
                   "Example Output for Batch Mode")


def load(name):
    return np.loadtxt(os.path.join(src, name))


def sig(a, digits=6):
    return [float(f"{x:.{digits}e}") for x in np.asarray(a, dtype=np.float64).ravel()]


x = load("rho_phi_EgridX.txt")
y = load("rho_phi_EgridY.txt")
de = load("density2d_e.txt")
di = load("density2d_i1.txt")
td = load("timeData.txt")
N = x.shape[0]
win = td[td[:, 0] >= 52000]
out = {
    "nodes": N,
    "graph_start": 52000,
    "graph_steps": 10400,
    "run_time": 62405,
    "centre_row": {"rho": sig(x[:, 1]), "phi": sig(x[:, 2]), "ex": sig(x[:, 3])},
    "centre_col": {"rho": sig(y[:, 1]), "phi": sig(y[:, 2]), "ey": sig(y[:, 3])},
    "density_e_centre_row": sig(de[N // 2]),
    "density_i_centre_row": sig(di[N // 2]),
    "density_e_total": float(de.sum()),
    "density_i_total": float(di.sum()),
    "time_rows": int(len(win)),
    "counts": [int(win[:, 1].min()), int(win[:, 1].max()), int(win[:, 2].min()), int(win[:, 2].max())],
    "probe_phi_mean": sig(win[:, 3:6].mean(axis=0)),
    "probe_phi_std": sig(win[:, 3:6].std(axis=0)),
    "probe_phi_series": sig(win[:, 3:6], 5),
}
dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "s17_stats.json")
with open(dst, "w") as f:
    json.dump(out, f, indent=0)
print("wrote", dst, os.path.getsize(dst), "bytes; probe means", out["probe_phi_mean"],
      "std", out["probe_phi_std"])
