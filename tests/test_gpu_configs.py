"""Parity at the benchmark workloads' own sizes.

C1 (workloads.plain: 65^2 nodes, 409 600 particles per species, physical
mode, mixed-precision multigrid to 1e-8): after 1 and 10 fused steps the GPU
tracks the FP64 reference-layout oracle (48-byte AoS, FP64 push and deposit,
the potential solved to a converged 1e-12) within 1e-6 relative L2 on
positions, velocities, phi and E; with the FP64 FMG solver it is bit-exact
against the build-layout oracle (cells, offsets, velocities, fixed-point
charge, phi, E, loss counts).

C2 (workloads.argon_256: 257^2 nodes, 13.1 M per species, argon MCC, pair
re-injection): three steps against the FP64 physical-mode MCC oracle drawing
the same Philox streams -- collision events per channel, wall losses and
particle counts agree to the few draws fp32 rounding moves across a decision
boundary; mean electron energy to 1e-5; charge is conserved exactly.

C3 (workloads.argon_1024: 1025^2 nodes, 104.9 M per species): two steps;
exact particle and charge bookkeeping, no overflow, and phi / E within 1e-6
of the oracle's converged FP64 solve of the GPU's own charge density.
"""
import numpy as np
import pytest

import orc
import pic2d
import workloads

pytestmark = pytest.mark.gpu

FX_ONE = 1 << orc.FX_TOTAL_BITS


def exact_sum(q):
    """sum of an int64 array without overflow (a node holds < 2^63, the
    grid's total does not fit 64 bits): high and low 32-bit halves apart"""
    q = np.asarray(q, dtype=np.int64).ravel()
    return (int((q >> 32).sum()) << 32) + int((q & 0xFFFFFFFF).sum())


def rel(a, b):
    a = np.asarray(a, dtype=np.float64).ravel()
    b = np.asarray(b, dtype=np.float64).ravel()
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def by_id(s):
    o = np.argsort(s["id"], kind="stable")
    return {k: v[o] for k, v in s.items()}


def fp64_physical_oracle(p, e0, i0, steps, mg_tol=1e-12):
    """the FP64 physical-mode step loop (orc_ref_run_mcc) with collisions off
    (no gas) and a converged potential"""
    N = p.nodes
    dx = p.length / (N - 1)
    _, cref = orc.configs_from_params(p)
    cref.mg_tol = mg_tol
    me, mi = orc.mcc_params(p, True), orc.mcc_params(p, False)
    for m in (me, mi):
        m.ng = 0.0
        m.nu_max = 0.0
        m.p_null = 0.0
    Ex = np.zeros((N + 1) * (N + 1))
    return orc.run_ref_mcc(cref, me, mi, False, p.vte, p.vti, steps, orc.state_to_aos(e0, dx),
                           orc.state_to_aos(i0, dx), e0["id"], i0["id"],
                           [len(e0["id"]), len(i0["id"])], Ex, Ex.copy(), grow=1.0)


def check_tracks_fp64(sim, rr, p, what, tol=1e-6):
    N = p.nodes
    dx = p.length / (N - 1)
    out = {}
    for s, key, kid in ((0, "e", "eid"), (1, "i", "iid")):
        g = by_id(sim.get_state(s))
        ref = rr[key][np.argsort(rr[kid], kind="stable")]
        rid = np.sort(rr[kid])
        common, ig, ir = np.intersect1d(g["id"], rid, assume_unique=True, return_indices=True)
        # a particle within rounding of a wall may leave on one side only
        assert len(common) >= 0.9999 * max(len(rid), len(g["id"])), (what, s, len(common))
        a = orc.state_to_aos({k: v[ig] for k, v in g.items()}, dx)
        b = ref[ir]
        for f in ("x", "y", "vx", "vy"):
            out[f"{key}.{f}"] = rel(a[f], b[f])
    gex, gey = sim.get_E()
    out["phi"] = rel(sim.get_phi(), rr["phi"])
    out["Ex"] = rel(gex[: N * N], rr["Ex"][: N * N])
    out["Ey"] = rel(gey[: N * N], rr["Ey"][: N * N])
    print(what, " ".join(f"{k} {v:.2e}" for k, v in out.items()), flush=True)
    for k, v in out.items():
        assert v < tol, (what, k, v)


def test_c1_full_size_tracks_fp64_oracle():
    p = workloads.plain()
    print(f"C1: {p.nodes}^2 nodes, {p.init_particle_no} per species, tol {p.mg_tol}", flush=True)
    with pic2d.Pic(p) as sim:
        e0, i0 = sim.get_state(0), sim.get_state(1)
        sim.run(1)  # the pre-loop wall pass (nobody is outside) and one step
        check_tracks_fp64(sim, fp64_physical_oracle(p, e0, i0, 1), p, "C1 step 1")
        sim.advance(9)
        check_tracks_fp64(sim, fp64_physical_oracle(p, e0, i0, 10), p, "C1 step 10")
        assert sim.overflow() == 0


def test_c1_full_size_bit_exact_vs_build_oracle():
    p = workloads.plain()
    p.mg_mode = pic2d.MG_FIXED
    with pic2d.Pic(p) as sim:
        e0, i0, E0 = sim.get_state(0), sim.get_state(1), sim.get_E32()
        c32, _ = orc.configs_from_params(p)
        c32.preloop = 1
        r = orc.run32(c32, 1, e0, i0, E0)
        sim.run(1)
        c32.preloop = 0
        c32.start_step = 1
        r9 = orc.run32(c32, 9, r["e"], r["i"], r["E32"])
        sim.advance(9)
        for s, key in ((0, "e"), (1, "i")):
            g, o = by_id(sim.get_state(s)), by_id(r9[key])
            assert np.array_equal(g["id"], o["id"]), key
            for k in ("cell", "ox", "oy", "vx", "vy", "vz"):
                assert np.array_equal(g[k], o[k]), (key, k)
        assert np.array_equal(sim.get_qfixed(0), r9["qe"])
        assert np.array_equal(sim.get_qfixed(1), r9["qi"])
        assert np.array_equal(sim.get_phi(), r9["phi"])
        assert np.array_equal(sim.get_E32(), r9["E32"])
        lost = np.concatenate([sim.loss_counts(0)[:4], sim.loss_counts(1)[:4]])
        assert np.array_equal(lost, r["lost"] + r9["lost"])
        # charge: every particle in the store deposits exactly 2^48
        for s in (0, 1):
            assert exact_sum(sim.get_qfixed(s)) == sim.particle_count(s) * FX_ONE


def mean_energy_ev(vx, vy, vz, mass):
    v2 = vx.astype(np.float64) ** 2 + vy.astype(np.float64) ** 2 + vz.astype(np.float64) ** 2
    return 0.5 * mass * v2.mean() / 1.602177e-19


def test_c2_full_size_against_fp64_mcc_oracle():
    p = workloads.argon_256()
    steps = 3
    print(f"C2: {p.nodes}^2 nodes, {p.init_particle_no} per species, {steps} steps", flush=True)
    N = p.nodes
    dx = p.length / (N - 1)
    with pic2d.Pic(p) as sim:
        e0, i0 = sim.get_state(0), sim.get_state(1)
        sim.run(steps)
        sim.sync()
        assert sim.overflow() == 0
        ev = [sim.collision_counts(s) for s in (0, 1)]
        lost = [sim.loss_counts(s) for s in (0, 1)]
        cnt = [sim.particle_count(s) for s in (0, 1)]
        qsum = [exact_sum(sim.get_qfixed(s)) for s in (0, 1)]
        ge = sim.get_state(0)
    # exact bookkeeping: an ion lost to a wall returns as a re-injected pair;
    # an ionisation adds one electron and one ion
    iz = ev[0][3]
    lost_i = int(lost[1][:4].sum())
    assert cnt[1] == p.init_particle_no + iz
    assert cnt[0] == p.init_particle_no - int(lost[0][:4].sum()) + iz + lost_i
    for s in (0, 1):
        assert qsum[s] == cnt[s] * FX_ONE
    _, cref = orc.configs_from_params(p)
    cref.mg_tol = 1e-10
    me, mi = orc.mcc_params(p, True), orc.mcc_params(p, False)
    Ex = np.zeros((N + 1) * (N + 1))
    rr = orc.run_ref_mcc(cref, me, mi, True, p.vte, p.vti, steps, orc.state_to_aos(e0, dx),
                         orc.state_to_aos(i0, dx), e0["id"], i0["id"],
                         [len(e0["id"]), len(i0["id"])], Ex, Ex.copy(), grow=1.05)
    oev = rr["events"]
    gev = ev[0] + ev[1]
    print("C2 events gpu", gev[:6].tolist(), "oracle", oev[:6].tolist(), flush=True)
    for k in (1, 2, 3, 4, 5):
        # the same draws: only a draw that fp32 rounding moves across a channel
        # boundary differs (measured <= 1 in 2e4 of the candidates)
        assert abs(int(gev[k]) - int(oev[k])) <= 5 + 2e-3 * oev[k], (k, gev[k], oev[k])
    assert oev[3] > 0 and oev[2] > 0 and oev[1] > 0 and oev[4] > 0
    ne_o, ni_o = len(rr["e"]), len(rr["i"])
    print(f"C2 counts gpu {cnt} oracle {[ne_o, ni_o]}", flush=True)
    assert abs(cnt[0] - ne_o) <= 5 + 1e-5 * ne_o and abs(cnt[1] - ni_o) <= 5 + 1e-5 * ni_o
    me_gpu = mean_energy_ev(ge["vx"], ge["vy"], ge["vz"], p.electron_mass)
    me_orc = mean_energy_ev(rr["e"]["vx"], rr["e"]["vy"], rr["e"]["vz"], p.electron_mass)
    print(f"C2 mean electron energy gpu {me_gpu:.6f} eV oracle {me_orc:.6f} eV", flush=True)
    assert abs(me_gpu / me_orc - 1) < 1e-5


def test_c3_full_size_bookkeeping_and_converged_potential():
    p = workloads.argon_1024()
    steps = 2
    N = p.nodes
    dx = p.length / (N - 1)
    print(f"C3: {N}^2 nodes, {p.init_particle_no} per species, {steps} steps", flush=True)
    with pic2d.Pic(p) as sim:
        sim.run(steps)
        sim.sync()
        assert sim.overflow() == 0
        ev = sim.collision_counts(0)
        lost = [sim.loss_counts(s) for s in (0, 1)]
        cnt = [sim.particle_count(s) for s in (0, 1)]
        iz = int(ev[3])
        lost_i = int(lost[1][:4].sum())
        assert cnt[1] == p.init_particle_no + iz
        assert cnt[0] == p.init_particle_no - int(lost[0][:4].sum()) + iz + lost_i
        for s in (0, 1):
            assert exact_sum(sim.get_qfixed(s)) == cnt[s] * FX_ONE
        rho = sim.get_rho()
        phi = sim.get_phi()
        gex, gey = sim.get_E()
        cyc, _, rr = sim.mg_stats()
    assert rr <= p.mg_tol
    # the oracle's FP64 multigrid, converged, on the same charge density
    f = -rho / p.epsilon
    u, _, rref = orc.mg_tol(f, p.length, 1e-12, max_cycles=200)
    assert rref <= 1e-12
    ex, ey = orc.ref_efield(u, dx)
    e_phi, e_ex, e_ey = rel(phi, u), rel(gex[: N * N], ex[: N * N]), rel(gey[: N * N], ey[: N * N])
    print(f"C3 phi err {e_phi:.2e} Ex {e_ex:.2e} Ey {e_ey:.2e} (tol {p.mg_tol}, {cyc} cycles)",
          flush=True)
    assert e_phi < 1e-6 and e_ex < 1e-6 and e_ey < 1e-6
