"""Collisional runs against the FP64 oracle: the parity target of the
collisional workloads.

The GPU timestep (fp32 state, cell + offset layout) and the oracle's FP64
reference-layout step loop (orc_ref_run_mcc) start from the same seeded
particles and draw collisions from the same Philox streams.  Their
trajectories drift apart through fp32 rounding, so the comparison is
statistical: ionisation count, mean electron energy and the time-averaged
density profiles must agree within 3 standard deviations of their sampling
error.  The argon background, 10 eV electrons (so that ionisation is
frequent) and pair re-injection follow the 256x256 argon workload at a size
the oracle runs in seconds.
"""
import math

import numpy as np
import pytest

import orc
import pic2d
import workloads

pytestmark = pytest.mark.gpu

Q, ME = 1.602177e-19, 9.10939e-31


def profile(state, N, nbins):
    """particles per x-bin (column index of the cell)"""
    jx, _ = pic2d.cell_unpack(state["cell"])
    x = (jx + state["ox"].astype(np.float64)) / (N - 1)
    return np.histogram(x, bins=nbins, range=(0.0, 1.0))[0].astype(np.float64)


def profile_aos(a, L, nbins):
    return np.histogram(a["x"] / L, bins=nbins, range=(0.0, 1.0))[0].astype(np.float64)


def energy_ev(vx, vy, vz):
    v2 = np.asarray(vx, np.float64) ** 2 + np.asarray(vy, np.float64) ** 2 + \
        np.asarray(vz, np.float64) ** 2
    return 0.5 * ME * v2 / Q


def within(a, b, sa, sb, z=3.0):
    return abs(a - b) <= z * math.sqrt(sa * sa + sb * sb) + 1e-300


@pytest.mark.parametrize("plop", [False, True])
def test_collisional_run_matches_fp64_oracle_statistically(plop):
    p = workloads.with_argon(workloads.discharge(32, 60, te=10.0, ti=0.5,
                                                 mg_mode=pic2d.MG_FIXED))
    if not plop:
        p.wall = pic2d.WALL_ABSORB
    steps, avg_from, nb = 40, 20, 16
    N, L = p.nodes, p.length
    with pic2d.Pic(p) as sim:
        e0, i0 = sim.get_state(0), sim.get_state(1)
        prof_e = np.zeros(nb)
        prof_i = np.zeros(nb)
        for s in range(steps):
            sim.advance(1)
            if s >= avg_from:
                prof_e += profile(sim.get_state(0), N, nb)
                prof_i += profile(sim.get_state(1), N, nb)
        ev_gpu = sim.collision_counts(0)
        e1 = sim.get_state(0)
        ni_gpu = sim.particle_size(1)
    _, cref = orc.configs_from_params(p)
    me, mi = orc.mcc_params(p, True), orc.mcc_params(p, False)
    ea, ia = orc.state_to_aos(e0, p.dx), orc.state_to_aos(i0, p.dx)
    Ex = np.zeros((N + 1) ** 2)
    Ey = np.zeros((N + 1) ** 2)
    eid, iid = e0["id"], i0["id"]
    nid = [len(ea), len(ia)]
    ev_orc = np.zeros(8, np.int64)
    oprof_e = np.zeros(nb)
    oprof_i = np.zeros(nb)
    for s in range(steps):
        cref.start_step = s
        r = orc.run_ref_mcc(cref, me, mi, plop and p.wall == pic2d.WALL_STEADY, p.vte, p.vti, 1,
                            ea, ia, eid, iid, nid, Ex, Ey)
        ea, ia, eid, iid, nid, Ex, Ey = r["e"], r["i"], r["eid"], r["iid"], r["next_id"], \
            r["Ex"], r["Ey"]
        ev_orc += r["events"]
        if s >= avg_from:
            oprof_e += profile_aos(ea, L, nb)
            oprof_i += profile_aos(ia, L, nb)
    # ionisation count: Poisson errors
    g, o = float(ev_gpu[3]), float(ev_orc[3])
    assert g > 500 and o > 500
    assert within(g, o, math.sqrt(g), math.sqrt(o)), (g, o)
    for k in (1, 2):
        g, o = float(ev_gpu[k]), float(ev_orc[k])
        assert within(g, o, math.sqrt(g), math.sqrt(o)), (k, g, o)
    # mean electron energy at the end
    Eg = energy_ev(e1["vx"], e1["vy"], e1["vz"])
    Eo = energy_ev(ea["vx"], ea["vy"], ea["vz"])
    assert within(Eg.mean(), Eo.mean(), Eg.std() / math.sqrt(len(Eg)),
                  Eo.std() / math.sqrt(len(Eo))), (Eg.mean(), Eo.mean())
    # particle counts and time-averaged density profiles
    assert within(len(Eg), len(Eo), math.sqrt(len(Eg) * 0.01), math.sqrt(len(Eo) * 0.01))
    if plop:
        assert ni_gpu == len(i0["id"]) + int(ev_gpu[3])   # every lost ion replaced
        assert len(ia) == len(i0["id"]) + int(ev_orc[3])
    nsamp = steps - avg_from
    for gp, op in ((prof_e, oprof_e), (prof_i, oprof_i)):
        # each bin's time average: the sampling error of one snapshot's count
        # (Poisson) is the conservative bound, snapshots being correlated
        sg = np.sqrt(gp / nsamp) * nsamp
        so = np.sqrt(op / nsamp) * nsamp
        bad = np.abs(gp - op) > 3.0 * np.sqrt(sg ** 2 + so ** 2) + 1
        assert not bad.any(), (gp / nsamp, op / nsamp)
