"""Particle-count control on the GPU against the oracle (orc_control.c).

* Source (particle.cpp:491-510): the injected particles of the phase entry
  point equal orc_inject's (cells and offsets bit for bit; velocities go
  through double cos/log/sqrt, so they are compared to 2e-7 relative); in the
  fused step, injected electrons are pushed and wall-tested with the step's
  field, injected ions join the ion pass, and the charge they add is their
  weighted deposit (weight 0 in reference mode, particleVec() value-init).
* Reduction (particle.cpp:520-653): one fused step with the reduction on
  equals the same step with it off followed by orc_reduce -- survivors,
  weights and the fixed-point charge bit for bit.
* Weights: a dataDump with mixed weights restarts with them, and the charge
  is orc32_deposit_w of the state.
"""
import numpy as np
import pytest

import orc
import pic2d
import workloads

pytestmark = pytest.mark.gpu

KEYS = ("cell", "ox", "oy", "vx", "vy", "vz")


def by_id(s):
    o = np.argsort(s["id"], kind="stable")
    return {k: v[o] for k, v in s.items()}


def make(mode, inject=0, reduce=0, cells=32, ppc=20, **kw):
    p = workloads.discharge(cells, ppc, mode=mode, mg_mode=pic2d.MG_FIXED)
    p.number_to_inject = inject
    p.reduce = reduce
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def close_vel(a, b):
    return np.all(np.abs(a - b) <= 2e-7 * np.maximum(np.abs(a), np.abs(b)) + 1e-30)


@pytest.mark.parametrize("mode", [pic2d.MODE_REF, pic2d.MODE_PHYSICAL])
def test_inject_phase_matches_oracle(mode):
    p = make(mode, inject=500)
    print(f"inject phase: mode={mode} N={p.nodes}", flush=True)
    with pic2d.Pic(p) as sim:
        for s, vt in ((0, p.vte), (1, p.vti)):
            n0 = sim.particle_size(s)
            sim.inject_particles(s)
            st = by_id(sim.get_state(s))
            w = sim.get_weights(s)
            assert len(st["id"]) == n0 + 500
            new = st["id"] >= n0
            got = {k: v[new] for k, v in st.items()}
            ref = orc.inject(500, s, 0, p.seed, p.dx, p.length, vt)
            for k in ("cell", "ox", "oy"):
                assert np.array_equal(got[k], ref[k]), k
            for k in ("vx", "vy", "vz"):
                assert close_vel(got[k], ref[k]), k
            order = np.argsort(sim.get_state(s)["id"], kind="stable")
            wn = w[order][new]
            assert np.all(wn == (0.0 if mode == pic2d.MODE_REF else 1.0))
            assert np.all(w[order][~new] == 1.0)
            assert sim.population_stats(s)[2] == 500


@pytest.mark.parametrize("mode", [pic2d.MODE_REF, pic2d.MODE_PHYSICAL])
def test_inject_fused_step_matches_oracle(mode):
    n_inj = 300
    pa = make(mode, inject=n_inj)
    pb = make(mode, inject=0)
    print(f"inject fused: mode={mode} N={pa.nodes}", flush=True)
    c32, _ = orc.configs_from_params(pa)
    with pic2d.Pic(pa) as a, pic2d.Pic(pb) as b:
        n0 = [a.particle_size(0), a.particle_size(1)]
        E0 = a.get_E32()
        a.run(1)
        b.run(1)
        wexp = 0.0 if mode == pic2d.MODE_REF else 1.0
        for s, vt, pp in ((0, pa.vte, c32.pe), (1, pa.vti, c32.pi)):
            sa, sb = by_id(a.get_state(s)), by_id(b.get_state(s))
            old = sa["id"] < n0[s]
            # the other particles are untouched by the source
            assert np.array_equal(sa["id"][old], sb["id"])
            for k in KEYS:
                assert np.array_equal(sa[k][old], sb[k]), (s, k)
            inj = orc.inject(n_inj, s, 0, pa.seed, pa.dx, pa.length, vt)
            moved, status = orc.push32(inj, E0, pp)
            kept = status == 0
            got = {k: v[~old] for k, v in sa.items()}
            assert len(got["id"]) == kept.sum()
            for k in ("cell",):
                assert np.array_equal(got[k], moved[k][kept]), (s, k)
            for k in ("ox", "oy", "vx", "vy", "vz"):
                assert close_vel(got[k], moved[k][kept]), (s, k)
            # their charge: the weighted deposit of what the GPU holds
            w = np.full(len(got["id"]), wexp, dtype=np.float32)
            dq = a.get_qfixed(s) - b.get_qfixed(s)
            assert np.array_equal(dq, orc.deposit32_w(got, w, pa.nodes, c32.mode))
        # population bookkeeping over more steps
        a.run(5)
        for s in (0, 1):
            st = a.population_stats(s)
            assert st[2] == 6 * n_inj
            lost = a.loss_counts(s)[:4].sum()
            assert a.particle_size(s) == n0[s] + st[2] - lost


def reduce_params(p):
    p.size_to_check[0] = 0
    p.size_to_check[1] = 0
    p.no_groups = 3
    for s, vt in ((0, p.vte), (1, p.vti)):
        for g, (vl, gl, gn) in enumerate(((2 * vt, 5000, 4000), (3 * vt, 1000, 800),
                                          (1000 * vt, 100, 50))):
            p.v_limit[s][g] = vl
            p.group_limit[s][g] = gl
            p.new_group_size[s][g] = gn
    return p


@pytest.mark.parametrize("mode", [pic2d.MODE_REF, pic2d.MODE_PHYSICAL])
def test_reduction_step_matches_oracle(mode):
    pa = reduce_params(make(mode, reduce=1, sort_interval=4))
    pb = make(mode, reduce=0, sort_interval=4)
    print(f"reduction: mode={mode} N={pa.nodes} n={pa.init_particle_no}", flush=True)
    c32, _ = orc.configs_from_params(pa)
    with pic2d.Pic(pa) as a, pic2d.Pic(pb) as b:
        a.run(1)
        b.run(1)
        for s in (0, 1):
            sb = by_id(b.get_state(s))
            rc = orc.reduce_params(pa, s)
            keep, w1, ng, nr, removed = orc.reduce(rc, 0, sb["vx"], sb["vy"], sb["id"],
                                                   np.ones(len(sb["id"]), np.float32))
            assert removed > 0 and nr[:3].all(), (s, ng, nr)
            order = np.argsort(a.get_state(s)["id"], kind="stable")
            sa = by_id(a.get_state(s))
            wa = a.get_weights(s)[order]
            assert np.array_equal(sa["id"], sb["id"][keep]), s
            for k in KEYS:
                assert np.array_equal(sa[k], sb[k][keep]), (s, k)
            assert np.array_equal(wa, w1[keep]), s
            exp_q = orc.deposit32_w({k: v[keep] for k, v in sb.items()}, w1[keep], pa.nodes,
                                    c32.mode)
            assert np.array_equal(a.get_qfixed(s), exp_q), s
            st = a.population_stats(s)
            assert st[0] == 1 and st[1] == removed
        # the next steps run on the holes the reduction left (in-place passes)
        a.run(3)
        for s in (0, 1):
            st = a.get_state(s)
            assert len(np.unique(st["id"])) == len(st["id"]) == a.particle_size(s)
            q = orc.deposit32_w(st, a.get_weights(s), pa.nodes, c32.mode)
            assert np.array_equal(a.get_qfixed(s), q)


def test_contain_particle_growth_phase():
    p = reduce_params(make(pic2d.MODE_PHYSICAL, reduce=1))
    with pic2d.Pic(p) as sim:
        st = by_id(sim.get_state(0))
        sim.set_step(3)
        sim.contain_particle_growth(0)
        rc = orc.reduce_params(p, 0)
        keep, w1, _, _, removed = orc.reduce(rc, 3, st["vx"], st["vy"], st["id"],
                                             np.ones(len(st["id"]), np.float32))
        got = by_id(sim.get_state(0))
        assert np.array_equal(got["id"], st["id"][keep])
        order = np.argsort(sim.get_state(0)["id"], kind="stable")
        assert np.array_equal(sim.get_weights(0)[order], w1[keep])
        assert sim.population_stats(0)[1] == removed


def test_weighted_dump_restart(tmp_path):
    p = make(pic2d.MODE_PHYSICAL, cells=32, ppc=5)
    c32, _ = orc.configs_from_params(p)
    with pic2d.Pic(p) as sim:
        sim.dump_data(tmp_path, 17)
    rng = np.random.default_rng(8)
    for name in ("dataDump_e.txt", "dataDump_i1.txt"):
        d = np.loadtxt(tmp_path / name)
        d[:, 5] = rng.uniform(0.25, 4.0, len(d)).astype(np.float32)
        with open(tmp_path / name, "w") as f:
            for r in d:
                f.write(" ".join("%.17g" % v for v in r) + "\n")
    with pic2d.Pic(p, restart=True, dump_dir=tmp_path) as r:
        for s, name in ((0, "dataDump_e.txt"), (1, "dataDump_i1.txt")):
            d = np.loadtxt(tmp_path / name)
            aos = r.get_particles(s)
            st = r.get_state(s)
            w = r.get_weights(s)
            # ids are the dump's line numbers
            assert np.array_equal(w, d[st["id"], 5].astype(np.float32))
            assert np.array_equal(aos["weight"], d[st["id"], 5].astype(np.float32))
            r.collect_charge(s)
            assert np.array_equal(r.get_qfixed(s), orc.deposit32_w(st, w, p.nodes, c32.mode))
        # set_weights round trip
        w = np.linspace(0.5, 1.5, r.particle_size(0)).astype(np.float32)
        r.set_weights(0, w)
        assert np.array_equal(r.get_weights(0), w)
