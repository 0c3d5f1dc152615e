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
