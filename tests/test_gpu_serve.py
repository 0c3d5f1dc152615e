"""Ensembles and the serving loop on the GPU (include/pic2d.h: pic_batch_advance,
pic_serve): a run's results are bit-identical alone and inside a batch; the
serving loop completes every request, respects arrival order and the memory
budget; pic_estimate_bytes bounds what pic_create allocates."""
import numpy as np
import pytest

import pic2d
import workloads

pytestmark = pytest.mark.gpu


def small_runs():
    runs = []
    for k, (cells, ppc, gas) in enumerate(((32, 40, pic2d.GAS_ARGON), (64, 20, pic2d.GAS_HYDROGEN),
                                           (32, 60, pic2d.GAS_HYDROGEN), (64, 30, pic2d.GAS_ARGON))):
        p = workloads.discharge(cells, ppc, te=3.0, seed=500 + k)
        workloads.with_gas(p, gas, 13.332 * (1 + k))
        runs.append(p)
    return runs


def fingerprint(sim):
    st = [sim.get_state(s) for s in (0, 1)]
    return (tuple(len(x["id"]) for x in st), sim.get_qfixed(0).tobytes(), sim.get_qfixed(1).tobytes(),
            sim.get_phi().tobytes(), tuple(int(np.sort(x["id"])[-1]) for x in st))


def test_batch_results_equal_runs_alone():
    runs = small_runs()
    print(f"batch determinism: {len(runs)} runs", flush=True)
    alone = []
    for p in runs:
        with pic2d.Pic(p) as sim:
            sim.run(0)
            sim.advance(12)
            alone.append(fingerprint(sim))
    sims = [pic2d.Pic(p) for p in runs]
    try:
        for s in sims:
            s.run(0)
        pic2d.batch_advance(sims, 12)
        for k, s in enumerate(sims):
            assert fingerprint(s) == alone[k], k
    finally:
        for s in sims:
            s.close()


def test_estimate_bounds_allocation():
    for p in small_runs() + [workloads.argon_256(), workloads.plain()]:
        with pic2d.Pic(p) as sim:
            got = sim.device_bytes()
        est = pic2d.estimate_bytes(p)
        assert got <= est <= 1.6 * got + (64 << 20), (p.nodes, p.init_particle_no, got, est)


def test_serving_loop():
    runs = small_runs()
    reqs = [(p, 20 + 5 * k, 0.05 * k) for k, p in enumerate(runs)]
    res, st = pic2d.serve(reqs, int(100e9), max_active=8)
    assert all(r["status"] == 0 for r in res)
    for r, (p, steps, t) in zip(res, reqs):
        assert r["admit_s"] >= t - 1e-6 and r["finish_s"] >= r["admit_s"]
        assert r["particle_steps"] >= steps * 0.5 * 2 * p.init_particle_no
        assert 0 < r["step_p50_ms"] <= r["step_p99_ms"]
    assert st["particle_steps"] == sum(r["particle_steps"] for r in res)
    assert st["max_concurrent"] >= 1 and st["throughput"] > 0
    # a budget that fits one run at a time serialises them in arrival order
    need = max(pic2d.estimate_bytes(p) for p in runs)
    res, st = pic2d.serve([(p, 10, 0.0) for p in runs], need, max_active=8)
    assert st["max_concurrent"] == 1 and st["peak_bytes"] <= need
    admits = [r["admit_s"] for r in res]
    finishes = [r["finish_s"] for r in res]
    assert all(admits[k + 1] >= finishes[k] - 1e-3 for k in range(len(runs) - 1))


# ---------------------------------------------------------------- the server
BUDGET = 200 * 10 ** 9


def serve_alone(p, steps):
    with pic2d.Server(BUDGET, 1) as srv:
        return srv.wait(srv.submit(p, steps), timeout=240)


def test_server_results_equal_runs_alone_and_report_phi_error():
    runs = small_runs()[:3]
    steps = 15
    alone = [serve_alone(p, steps) for p in runs]
    for st in alone:
        assert st["state"] == pic2d.REQ_DONE and st["error"] == 0
        assert 0 <= st["phi_error"] < 1e-9 and 0 <= st["e_error"] < 1e-8, st
    with pic2d.Server(BUDGET, 8) as srv:
        ids = [srv.submit(p, steps) for p in runs]
        mixed = [srv.wait(i, timeout=240) for i in ids]
    print("fingerprints alone", [hex(s["fingerprint"]) for s in alone],
          "mixed", [hex(s["fingerprint"]) for s in mixed], flush=True)
    for a, m in zip(alone, mixed):
        assert m["state"] == pic2d.REQ_DONE and m["steps_done"] == steps
        assert m["fingerprint"] == a["fingerprint"] and m["particle_steps"] == a["particle_steps"]
        assert m["phi_error"] == a["phi_error"]


def test_server_quarantines_a_non_finite_run():
    """a run whose charge scale is NaN gets a NaN potential on its first step:
    it stops with PIC_ERR_NUMERIC and frees its slot; the runs beside it finish
    with the results they have alone"""
    runs = small_runs()[:2]
    steps = 12
    alone = [serve_alone(p, steps) for p in runs]
    bad = small_runs()[2]
    bad.qce = float("nan")
    bad.qci = float("nan")
    with pic2d.Server(BUDGET, 8) as srv:
        ids = [srv.submit(runs[0], steps), srv.submit(bad, steps), srv.submit(runs[1], steps)]
        st = [srv.wait(i, timeout=240) for i in ids]
    assert st[1]["state"] == pic2d.REQ_FAILED and st[1]["error"] == pic2d.PIC_ERR_NUMERIC
    assert st[1]["steps_done"] == 1
    for a, m in zip(alone, (st[0], st[2])):
        assert m["state"] == pic2d.REQ_DONE and m["fingerprint"] == a["fingerprint"]


def test_server_cancel_mid_run_and_fifo_admission():
    runs = small_runs()
    long_run, short = runs[3], runs[0]
    want = serve_alone(short, 10)
    with pic2d.Server(BUDGET, 2) as srv:
        lid = srv.submit(long_run, 100000)
        sid = srv.submit(short, 10)
        st = srv.poll(lid)
        while st["steps_done"] < 20 and st["state"] in (pic2d.REQ_QUEUED, pic2d.REQ_RUNNING):
            st = srv.poll(lid)
        srv.cancel(lid)
        ls = srv.wait(lid, timeout=120)
        ss = srv.wait(sid, timeout=240)
        assert ls["state"] == pic2d.REQ_CANCELLED and 20 <= ls["steps_done"] < 100000
        assert ss["state"] == pic2d.REQ_DONE and ss["fingerprint"] == want["fingerprint"]
    # one slot: the second request is admitted after the first finishes
    with pic2d.Server(BUDGET, 1) as srv:
        a, b = srv.submit(short, 5), srv.submit(runs[1], 5)
        sa, sb = srv.wait(a, timeout=240), srv.wait(b, timeout=240)
        assert sb["admit_s"] >= sa["finish_s"] and sa["admit_s"] <= sb["admit_s"]
        # a queued request cancelled before its turn never runs
        c, d = srv.submit(long_run, 100000), srv.submit(short, 5)
        srv.cancel(d)
        sd = srv.wait(d, timeout=60)
        assert sd["state"] == pic2d.REQ_CANCELLED and sd["steps_done"] == 0
        srv.cancel(c)
        assert srv.wait(c, timeout=120)["state"] == pic2d.REQ_CANCELLED
