"""Mixed-precision multigrid (FP32 V-cycles, FP64 residual and accumulation)
against the oracle's FP64 multigrid converged to 1e-13: phi within 1e-9
relative L2 once the residual target is 1e-11; inside a full timestep the
potential matches the oracle's converged potential for the same (bit-exact)
charge."""
import numpy as np
import pytest

import orc
import pic2d
import workloads

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [33, 65, 129, 257, 513, 1025])
def test_poisson_tol_matches_converged_fp64(n):
    print(f"poisson_tol n={n}", flush=True)
    rng = np.random.default_rng(n)
    L = 0.01
    f = rng.standard_normal((n, n)) * 1e9
    u, cycles, rr = pic2d.poisson_tol(f, length=L, tol=1e-11, max_iter=80)
    assert rr <= 1e-11 and 1 <= cycles <= 80
    ref, _, rref = orc.mg_tol(f, L, 1e-13, max_cycles=300)
    assert rref <= 1e-13
    err = np.linalg.norm(u - ref) / np.linalg.norm(ref)
    assert err < 1e-9, err
    # the residual the GPU reports is the true FP64 residual of its iterate
    assert abs(orc.relres(u, f, L) - rr) <= 1e-3 * rr + 1e-15


def test_poisson_tol_eigenmode_and_warm_start():
    n, L = 257, 0.02
    h = L / (n - 1)
    x = np.arange(n) * h
    f = np.outer(np.sin(2 * np.pi * x / L), np.sin(3 * np.pi * x / L))
    lam = -(4 / h ** 2) * (np.sin(2 * np.pi * h / (2 * L)) ** 2 + np.sin(3 * np.pi * h / (2 * L)) ** 2)
    exact = f / lam
    u, c1, _ = pic2d.poisson_tol(f, length=L, tol=1e-12)
    assert np.max(np.abs(u - exact)) < 1e-9 * np.max(np.abs(exact))
    # warm start from the solution: already converged, no cycle needed
    u2, c2, rr2 = pic2d.poisson_tol(f, u0=u, length=L, tol=1e-10)
    assert c2 == 0 and rr2 <= 1e-10


@pytest.mark.parametrize("sweeps", [1, 2, 3])
def test_timestep_potential_in_tolerance_mode(sweeps):
    """V(1,1) .. V(3,3) cycles reach the same converged potential"""
    print(f"timestep tol sweeps={sweeps}", flush=True)
    p = workloads.discharge(64, 20, mode=pic2d.MODE_PHYSICAL, mg_mode=pic2d.MG_TOL)
    p.mg_tol = 1e-11
    p.mg_max_iter = 60
    p.mg_sweeps = sweeps
    with pic2d.Pic(p) as sim:
        e0, i0, E0 = sim.get_state(0), sim.get_state(1), sim.get_E32()
        sim.run(1)
        cyc, _, rr = sim.mg_stats()
        assert rr <= 1e-11 and cyc >= 1
        phi = sim.get_phi()
        qe = sim.get_qfixed(0)
    c32, _ = orc.configs_from_params(p)
    c32.preloop = 1
    c32.mg_mode = 1
    c32.mg_tol = 1e-13
    c32.mg_max_cycles = 300
    r = orc.run32(c32, 1, e0, i0, E0)
    assert np.array_equal(r["qe"], qe)            # same charge, bit for bit
    err = np.linalg.norm(phi - r["phi"]) / np.linalg.norm(r["phi"])
    assert err < 1e-9, err


def test_more_sweeps_need_fewer_cycles():
    """the smoothing per leg trades against the cycle count: from the same
    start, V(2,2) reaches 1e-10 in fewer cycles than V(1,1)"""
    cycles = {}
    for nu in (1, 2):
        p = workloads.discharge(128, 10, mode=pic2d.MODE_PHYSICAL, mg_mode=pic2d.MG_TOL)
        p.mg_tol = 1e-10
        p.mg_max_iter = 60
        p.mg_sweeps = nu
        with pic2d.Pic(p) as sim:
            sim.run(1)
            cycles[nu], _, rr = sim.mg_stats()
            assert rr <= 1e-10
    assert cycles[2] < cycles[1], cycles


@pytest.mark.parametrize("tol", [1e-8, 1e-10])
def test_tolerance_mode_steps_against_converged_fp64(tol):
    """12 fused steps at a residual target (1e-8 is the benchmark's): from the
    second step the solve starts from the linear extrapolation of the last
    two potentials, the device picks the final iterate by the applied cycle
    count's parity and the host swaps the buffers across steps.  After every
    step phi and E are compared with the oracle's FP64 multigrid converged to
    1e-13 on the GPU's own (bit-exact fixed-point) charge density."""
    p = workloads.discharge(64, 20, mode=pic2d.MODE_PHYSICAL, mg_mode=pic2d.MG_TOL, te=4.0)
    p.mg_tol = tol
    N = p.nodes
    dx = p.length / (N - 1)
    worst = [0.0, 0.0]
    with pic2d.Pic(p) as sim:
        sim.run(1)
        for step in range(12):
            if step:
                sim.advance(1)
            cyc, _, rr = sim.mg_stats()
            assert rr <= tol, (step, rr)
            rho, phi = sim.get_rho(), sim.get_phi()
            gex, gey = sim.get_E()
            u, _, rref = orc.mg_tol(-rho / p.epsilon, p.length, 1e-13, max_cycles=300)
            assert rref <= 1e-13
            ex, ey = orc.ref_efield(u, dx)
            e_phi = np.linalg.norm(phi - u) / np.linalg.norm(u)
            e_E = max(np.linalg.norm(gex - ex) / np.linalg.norm(ex),
                      np.linalg.norm(gey - ey) / np.linalg.norm(ey))
            worst = [max(worst[0], e_phi), max(worst[1], e_E)]
            # E32 (the gather's copy) is the fp32 rounding of the FP64 E
            E32 = sim.get_E32()
            k = (N + 1) * N + 1
            assert np.array_equal(E32[0::2], gex[:k].astype(np.float32))
            assert np.array_equal(E32[1::2], gey[:k].astype(np.float32))
    print(f"tol {tol}: worst phi err {worst[0]:.2e}, E err {worst[1]:.2e} over 12 steps",
          flush=True)
    assert worst[0] < 1e-6 and worst[1] < 1e-6
    if tol <= 1e-10:
        assert worst[0] < 1e-9


def test_field_error_report_against_oracle():
    """pic_field_error: the library's FP64 reference solve reproduces the
    error the oracle measures.  With the reference's fixed-cycle FMG
    (ncycle 2, not converged) phi is off by a visible amount, which both
    must agree on; in tolerance mode the periodic report keeps one row per
    `every` steps with errors below the solver's target."""
    p = workloads.discharge(64, 20, mode=pic2d.MODE_PHYSICAL, mg_mode=pic2d.MG_FIXED, te=4.0)
    N = p.nodes
    dx = p.length / (N - 1)
    with pic2d.Pic(p) as sim:
        sim.run(3)
        fe = sim.field_error(1e-13)
        rho, phi = sim.get_rho(), sim.get_phi()
        gex, gey = sim.get_E()
    u, _, _ = orc.mg_tol(-rho / p.epsilon, p.length, 1e-13, max_cycles=300)
    ex, ey = orc.ref_efield(u, dx)
    want = (np.linalg.norm(phi - u) / np.linalg.norm(u),
            np.linalg.norm(gex - ex) / np.linalg.norm(ex),
            np.linalg.norm(gey - ey) / np.linalg.norm(ey))
    print("field_error fixed mode:", fe, "oracle:", want, flush=True)
    assert fe["ref_relres"] <= 1e-13
    assert want[0] > 1e-7  # fixed cycles leave a measurable error
    for got, w in zip((fe["phi"], fe["ex"], fe["ey"]), want):
        assert abs(got - w) <= 1e-4 * w
    p.mg_mode = pic2d.MG_TOL
    p.mg_tol = 1e-10
    with pic2d.Pic(p) as sim:
        sim.set_accuracy_report(3)
        sim.run(9)
        rows = sim.accuracy_report()
    print("accuracy report rows:", rows.tolist(), flush=True)
    assert rows.shape == (3, 6)
    assert np.array_equal(rows[:, 0], [2, 5, 8])
    assert np.all(rows[:, 1] <= 1e-10) and np.all(rows[:, 5] <= 1e-12)
    assert np.all(rows[:, 2:5] < 1e-8)


def test_graph_replayed_cycles_match_launched_cycles():
    """the tolerance solve's V-cycle batches replayed from captured hipGraphs
    give the states, charge, phi and E of the kernel-by-kernel launches, bit
    for bit, over steps that reuse the cached graphs"""
    out = []
    for graphs in (0, 1):
        p = workloads.discharge(128, 20, te=4.0)
        p.mg_tol = 1e-10
        p.graphs = graphs
        with pic2d.Pic(p) as sim:
            sim.run(12)
            out.append((sim.get_state(0), sim.get_qfixed(0), sim.get_phi(), sim.get_E32(),
                        sim.mg_stats()))
    a, b = out
    o0, o1 = np.argsort(a[0]["id"]), np.argsort(b[0]["id"])
    for k in ("id", "cell", "ox", "oy", "vx", "vy", "vz"):
        assert np.array_equal(a[0][k][o0], b[0][k][o1]), k
    assert np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2]) and np.array_equal(a[3], b[3])
    assert a[4][1] == b[4][1]  # the same number of V-cycles


def test_small_grid_verdicts_read_late_match_step_by_step():
    """grids up to 65^2 are solved by one workgroup whose verdict the host
    reads only before the next solve or a report: ten steps queued at once
    report the cycle total and final residual of ten steps each followed by
    a read, and reach the same state"""
    out = []
    for per_step in (False, True):
        p = workloads.discharge(64, 20, te=4.0)
        p.mg_tol = 1e-10
        with pic2d.Pic(p) as sim:
            last = []
            if per_step:
                for _ in range(10):
                    sim.advance(1)
                    last.append(sim.mg_stats()[0])
            else:
                sim.advance(10)
            stats = sim.mg_stats()
            out.append((sim.get_state(0), sim.get_phi(), stats, last))
    (s0, phi0, st0, _), (s1, phi1, st1, last) = out
    print("late:", st0, "per step:", st1, last, flush=True)
    assert st0 == st1
    assert sum(last) == st1[1] and all(c >= 1 for c in last) and last[-1] == st1[0]
    assert st1[2] <= 1e-10
    # the store order may differ (a call's end closes its fused sort
    # histogram), the particles and the charge may not
    o0, o1 = np.argsort(s0["id"]), np.argsort(s1["id"])
    for k in ("id", "cell", "ox", "oy", "vx", "vy", "vz"):
        assert np.array_equal(s0[k][o0], s1[k][o1]), k
    assert np.array_equal(phi0, phi1)
