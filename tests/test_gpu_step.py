"""Fused GPU timestep (pic_run) against the oracle.

Bit-exact against the build-layout restatement (orc32_run): particle cells,
offsets, velocities (matched by id), fixed-point charge, loss counts, phi and
the packed E, after 1 and 10 steps.  Within 1e-6 relative L2 against the FP64
reference-layout restatement (orc_ref_run) for positions, velocities, phi, E.
The per-phase entry points reproduce the fused step bit for bit.
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


def assert_states_equal(a, b, what):
    a, b = by_id(a), by_id(b)
    assert np.array_equal(a["id"], b["id"]), f"{what}: particle sets differ"
    for k in KEYS:
        bad = np.flatnonzero(a[k] != b[k])
        assert bad.size == 0, f"{what}.{k}: {bad.size} mismatches, first id {a['id'][bad[0]]}"


def make(cells=64, ppc=20, mode=pic2d.MODE_REF, **kw):
    p = workloads.discharge(cells, ppc, mode=mode, mg_mode=pic2d.MG_FIXED, **kw)
    return p, pic2d.Pic(p)


@pytest.mark.parametrize("mode", [pic2d.MODE_REF, pic2d.MODE_PHYSICAL])
def test_fused_step_bit_exact_vs_build_oracle(mode):
    p, sim = make(mode=mode)
    print(f"fused step parity: mode={mode} N={p.nodes} n={p.init_particle_no}", flush=True)
    with sim:
        e0, i0, E0 = sim.get_state(0), sim.get_state(1), sim.get_E32()
        c32, _ = orc.configs_from_params(p)
        c32.preloop = 1
        r = orc.run32(c32, 1, e0, i0, E0)
        sim.run(1)
        assert_states_equal(sim.get_state(0), r["e"], "e@1")
        assert_states_equal(sim.get_state(1), r["i"], "i@1")
        assert np.array_equal(sim.get_qfixed(0), r["qe"])
        assert np.array_equal(sim.get_qfixed(1), r["qi"])
        assert np.array_equal(sim.get_phi(), r["phi"])
        assert np.array_equal(sim.get_E32(), r["E32"])
        # nine more steps from the GPU's own state
        c32.preloop = 0
        c32.start_step = 1
        r9 = orc.run32(c32, 9, r["e"], r["i"], r["E32"])
        sim.run(9)
        assert_states_equal(sim.get_state(0), r9["e"], "e@10")
        assert_states_equal(sim.get_state(1), r9["i"], "i@10")
        assert np.array_equal(sim.get_qfixed(0), r9["qe"])
        assert np.array_equal(sim.get_phi(), r9["phi"])
        assert np.array_equal(sim.get_E32(), r9["E32"])
        lost = sim.loss_counts(0).sum() + sim.loss_counts(1).sum()
        assert lost == r["lost"].sum() + r9["lost"].sum()
        assert np.array_equal(np.concatenate([sim.loss_counts(0)[:4], sim.loss_counts(1)[:4]]),
                              r["lost"] + r9["lost"])
        assert sim.get_step() == 10


def test_fused_step_tracks_fp64_reference():
    p, sim = make(cells=64, ppc=20)
    N, dx = p.nodes, p.dx
    with sim:
        e0, i0 = sim.get_state(0), sim.get_state(1)
        _, cref = orc.configs_from_params(p)
        Ex = np.zeros((N + 1) * (N + 1))
        for steps in (1, 10):
            rr = orc.run_ref(cref, steps, orc.state_to_aos(e0, dx), orc.state_to_aos(i0, dx),
                             e0["id"], i0["id"], Ex, Ex.copy())
            with pic2d.Pic(p) as s2:
                s2.run(steps)
                g = by_id(s2.get_state(0))
                phi = s2.get_phi()
                gex, gey = s2.get_E()
            a = orc.state_to_aos(g, dx)
            idx = {int(k): j for j, k in enumerate(g["id"])}
            pairs = [(idx[int(k)], j) for j, k in enumerate(rr["eid"]) if int(k) in idx]
            assert len(pairs) >= 0.9999 * len(rr["eid"])
            ga, gb = a[[x for x, _ in pairs]], rr["e"][[y for _, y in pairs]]
            errs = {}
            for f in ("x", "y", "vx", "vy"):
                errs[f] = np.linalg.norm(ga[f] - gb[f]) / np.linalg.norm(gb[f])
            errs["phi"] = np.linalg.norm(phi - rr["phi"]) / np.linalg.norm(rr["phi"])
            errs["Ex"] = np.linalg.norm(gex - rr["Ex"]) / np.linalg.norm(rr["Ex"])
            errs["Ey"] = np.linalg.norm(gey - rr["Ey"]) / np.linalg.norm(rr["Ey"])
            print(f"{steps} steps vs FP64:", " ".join(f"{k} {v:.2e}" for k, v in errs.items()),
                  flush=True)
            for k, v in errs.items():
                assert v < 1e-6, (steps, k, v)


def test_phased_entry_points_reproduce_fused_step():
    p, fused = make(cells=32, ppc=30)
    _, phased = make(cells=32, ppc=30)
    with fused, phased:
        fused.run(3)
        for step in range(3):
            if step == 0:                      # pic.cpp:1049-1057 pre-loop wall pass
                phased.boundary(0)
                phased.boundary(1)
            phased.e_collisions()
            phased.get_particle_electric_field(0)
            phased.move_particles(0)
            phased.ion_collisions()
            phased.get_particle_electric_field(1)
            phased.move_particles(1)
            phased.boundary(0)
            phased.boundary(1)
            phased.collect_charge(1)
            phased.collect_charge(0)
            phased.grid_rho()
            phased.grid_potential(p.phi_left, p.phi_right, p.phi_bottom, p.phi_top, step,
                                  p.phi_period)
            phased.grid_electric_field()
        assert_states_equal(phased.get_state(0), fused.get_state(0), "e")
        assert_states_equal(phased.get_state(1), fused.get_state(1), "i")
        assert np.array_equal(phased.get_qfixed(0), fused.get_qfixed(0))
        assert np.array_equal(phased.get_rho(), fused.get_rho())
        assert np.array_equal(phased.get_phi(), fused.get_phi())
        a, b = phased.get_E(), fused.get_E()
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_loader_statistics_and_ids():
    p, sim = make(cells=64, ppc=50)
    with sim:
        for s, sig in ((0, p.sigma_load_e), (1, p.sigma_load_i)):
            st = sim.get_state(s)
            n = len(st["id"])
            assert n == p.init_particle_no
            assert np.array_equal(np.sort(st["id"]), np.arange(n, dtype=np.uint32))
            jx, jy = pic2d.cell_unpack(st["cell"])
            assert jx.min() >= 0 and jx.max() <= p.nodes - 2
            assert np.all((st["ox"] >= 0) & (st["ox"] < 1))
            for k in ("vx", "vy", "vz"):
                v = st[k].astype(np.float64)
                assert abs(v.mean()) < 5 * sig / np.sqrt(n)
                assert abs(v.std() / sig - 1) < 5 / np.sqrt(2 * n) + 1e-3
            # positions uniform: counts per cell within 6 sigma of the mean
            h = np.bincount(jy * (p.nodes - 1) + jx, minlength=(p.nodes - 1) ** 2)
            assert np.all(np.abs(h - h.mean()) < 6 * np.sqrt(h.mean()))
            # sorted by the counting-sort key after creation (8x8-cell tiles,
            # the library's default granularity)
            key = (((jy + 1) >> 3) * ((p.nodes + 8) // 8) + ((jx + 1) >> 3)) * 64 + \
                ((jy + 1) & 7) * 8 + ((jx + 1) & 7)
            assert np.all(np.diff(key >> 6) >= 0)


@pytest.mark.parametrize("sort_key", [0, 3, 6])
def test_sort_key_granularity(sort_key):
    """the sort orders the store by cell key >> sort_key; results do not depend on it"""
    p = workloads.discharge(64, 20, mode=pic2d.MODE_PHYSICAL, mg_mode=pic2d.MG_FIXED, te=6.0)
    p.sort_key = sort_key
    p.sort_interval = 1
    with pic2d.Pic(p) as sim:
        sim.run(3)
        st = sim.get_state(0)
        jx, jy = pic2d.cell_unpack(st["cell"])
        key = (((jy + 1) >> 3) * ((p.nodes + 8) // 8) + ((jx + 1) >> 3)) * 64 + \
            ((jy + 1) & 7) * 8 + ((jx + 1) & 7)
        assert np.all(np.diff(key >> sort_key) >= 0)
        got = by_id(st)
        q = sim.get_qfixed(0)
    p.sort_key = 0
    with pic2d.Pic(p) as ref:
        ref.run(3)
        want = by_id(ref.get_state(0))
        assert np.array_equal(q, ref.get_qfixed(0))
    for k in ("id",) + KEYS:
        assert np.array_equal(got[k], want[k]), k


def test_dump_and_restart_roundtrip(tmp_path):
    p, sim = make(cells=32, ppc=10)
    with sim:
        sim.run(2)
        sim.dump_data(tmp_path, 17)
        e, i = by_id(sim.get_state(0)), by_id(sim.get_state(1))
        ne, ni = len(e["id"]), len(i["id"])
    lines = open(tmp_path / "dataDump_e.txt").read().split("\n")
    assert len(lines) == ne + 1 and len(lines[0].split()) == 6
    with pic2d.Pic(p, restart=True, dump_dir=tmp_path) as r:
        assert r.particle_size(0) == ne and r.particle_size(1) == ni
        aos = r.get_particles(0)
    # restart positions are the dumped doubles to the last bit (17 digits)
    dumped = np.loadtxt(tmp_path / "dataDump_e.txt")
    assert np.array_equal(np.sort(aos["x"]), np.sort(dumped[:, 0]))
    assert np.array_equal(np.sort(aos["vx"]), np.sort(dumped[:, 2]))
    # truncated file is rejected (the reference would silently drop the record)
    txt = open(tmp_path / "dataDump_i1.txt").read()
    open(tmp_path / "dataDump_i1.txt", "w").write(txt[: len(txt) // 2] + " 1.0 garbage\n")
    with pytest.raises(pic2d.PicError) as err:
        pic2d.Pic(p, restart=True, dump_dir=tmp_path)
    assert err.value.code == -4


@pytest.mark.parametrize("mode", [pic2d.MODE_REF, pic2d.MODE_PHYSICAL])
def test_sort_interval_does_not_change_results(mode):
    """In-place passes between sort passes update the same particles with the
    same arithmetic; only the order in the store differs."""
    out = {}
    for k in (1, 4, 7):
        p = workloads.discharge(64, 20, mode=mode, mg_mode=pic2d.MG_FIXED, te=6.0)
        p.sort_interval = k
        print(f"sort interval {k} mode {mode}", flush=True)
        with pic2d.Pic(p) as sim:
            sim.run(15)
            out[k] = dict(e=sim.get_state(0), i=sim.get_state(1), qe=sim.get_qfixed(0),
                          qi=sim.get_qfixed(1), phi=sim.get_phi(), le=sim.loss_counts(0),
                          li=sim.loss_counts(1))
    for k in (4, 7):
        assert_states_equal(out[k]["e"], out[1]["e"], f"e k={k}")
        assert_states_equal(out[k]["i"], out[1]["i"], f"i k={k}")
        for key in ("qe", "qi", "phi", "le", "li"):
            assert np.array_equal(out[k][key], out[1][key]), (k, key)
    assert out[1]["le"][:4].sum() > 0


@pytest.mark.parametrize("sort_interval", [0, 1])
def test_reference_steady_walls_match_the_oracle(sort_interval):
    """REF mode with the steady-state walls: nobody leaves; each step the
    outside electrons and ions are paired per side in id order and
    re-injected.  Cells, offsets, ids and charge are bit-exact against the
    oracle after each step run from the same state; the re-injected
    velocities (fp32 cos/log of the device and of the host libm) agree to
    a few ulp.  With a sort pass every step and hot electrons, the unpaired
    particles that stay outside, some more than a cell beyond a wall, are
    sorted too (their sort keys are clamped to the tile grid)."""
    te = 8.0 if sort_interval == 0 else 80.0
    p = workloads.discharge(32, 40, mode=pic2d.MODE_REF, mg_mode=pic2d.MG_FIXED, te=te)
    p.wall = pic2d.WALL_STEADY
    p.sort_interval = sort_interval
    print(f"steady walls, sort interval {sort_interval}", flush=True)
    p.sigma_load_i = 3.0e5   # hot ions: ions leave too, so pairs form on every side
    with pic2d.Pic(p) as sim:
        c32, _ = orc.configs_from_params(p)
        assert c32.mode == orc.MODE_REF_STEADY
        pairs = 0
        for step in range(3):
            e0, i0, E0 = sim.get_state(0), sim.get_state(1), sim.get_E32()
            c32.preloop = 1 if step == 0 else 0
            c32.start_step = step
            r = orc.run32(c32, 1, e0, i0, E0)
            if step == 0:
                sim.run(1)
            else:
                sim.advance(1)
            e1, i1 = by_id(sim.get_state(0)), by_id(sim.get_state(1))
            re, ri = by_id(r["e"]), by_id(r["i"])
            for a, b, what in ((e1, re, "e"), (i1, ri, "i")):
                assert np.array_equal(a["id"], b["id"]), what
                assert len(a["id"]) == p.init_particle_no          # counts are constant
                for k in ("cell", "ox", "oy"):
                    assert np.array_equal(a[k], b[k]), (what, k, step)
                for k in ("vx", "vy", "vz"):
                    np.testing.assert_allclose(a[k], b[k], rtol=4e-7, atol=0,
                                               err_msg=f"{what}.{k} step {step}")
            assert np.array_equal(sim.get_qfixed(0), r["qe"])
            assert np.array_equal(sim.get_qfixed(1), r["qi"])
            moved = (i1["cell"] != by_id(i0)["cell"]) & (np.abs(i1["vx"]) > 0)
            pairs += int(moved.sum())
        assert pairs > 0
        assert sim.loss_counts(0).sum() == 0 and sim.loss_counts(1).sum() == 0
        far = 0
        for st in (sim.get_state(0), sim.get_state(1)):
            jx, jy = pic2d.cell_unpack(st["cell"])
            far += int(((jx < -1) | (jx > p.nodes) | (jy < -1) | (jy > p.nodes)).sum())
        print(f"particles more than a cell outside: {far}", flush=True)
        if sort_interval == 1:
            assert far > 0


@pytest.mark.parametrize("skew", [False, True])
def test_reference_steady_walls_many_candidates(skew):
    """The steady walls with tens of thousands of particles outside (45 % of
    the electrons, 30 % of the ions, on every side and in the corners, up to
    two cells beyond a wall): the pairing is bit-exact against the oracle.
    With `skew` every outside particle has an id below 2^16 and every inside
    one an id above 2^30, so all candidates of a side fall into one id bucket
    and the bucket is ordered by counting ranks instead of in LDS."""
    p = workloads.discharge(64, 20, mode=pic2d.MODE_REF, mg_mode=pic2d.MG_FIXED)
    p.wall = pic2d.WALL_STEADY
    N = p.nodes
    print(f"steady walls, many candidates, skew {skew}", flush=True)
    rng = np.random.default_rng(11)
    with pic2d.Pic(p) as sim:
        sim.run(1)
        states = []
        for s, frac in ((0, 0.45), (1, 0.3)):
            d = sim.get_state(s)
            n = len(d["id"])
            out = rng.random(n) < frac
            side = rng.integers(0, 6, n)  # left, right, bottom, top, left-bottom, right-top
            jx, jy = pic2d.cell_unpack(d["cell"])
            jx, jy = jx.astype(np.int64), jy.astype(np.int64)
            lo = -1 - rng.integers(0, 2, n)
            hi = N + rng.integers(0, 2, n)
            jx = np.where(out & np.isin(side, (0, 4)), lo, jx)
            jx = np.where(out & np.isin(side, (1, 5)), hi, jx)
            jy = np.where(out & np.isin(side, (2, 4)), lo, jy)
            jy = np.where(out & np.isin(side, (3, 5)), hi, jy)
            d["cell"] = pic2d.cell_pack(jx, jy)
            if skew:
                ids = np.empty(n, dtype=np.uint32)
                ids[out] = rng.permutation(int(out.sum())).astype(np.uint32)
                ids[~out] = (1 << 30) + rng.permutation(int((~out).sum())).astype(np.uint32)
                d["id"] = ids
            sim.set_state(s, d)
            states.append(d)
            print(f"species {s}: {int(out.sum())} outside of {n}", flush=True)
        E0 = sim.get_E32()
        c32, _ = orc.configs_from_params(p)
        c32.preloop = 0
        c32.start_step = 1
        r = orc.run32(c32, 1, states[0], states[1], E0)
        sim.advance(1)
        e1, i1 = by_id(sim.get_state(0)), by_id(sim.get_state(1))
        re, ri = by_id(r["e"]), by_id(r["i"])
        for a, b, what in ((e1, re, "e"), (i1, ri, "i")):
            assert np.array_equal(a["id"], b["id"]), what
            for k in ("cell", "ox", "oy"):
                bad = np.flatnonzero(a[k] != b[k])
                assert bad.size == 0, (what, k, bad.size)
            for k in ("vx", "vy", "vz"):
                np.testing.assert_allclose(a[k], b[k], rtol=4e-7, atol=0, err_msg=f"{what}.{k}")
        assert np.array_equal(sim.get_qfixed(0), r["qe"])
        assert np.array_equal(sim.get_qfixed(1), r["qi"])
        # most outside ions found a partner: well over 16384 pairs
        def outside(c):
            jx, jy = pic2d.cell_unpack(c)
            return (jx < 0) | (jx >= N) | (jy < 0) | (jy >= N)
        before = outside(by_id(states[1])["cell"])
        after = outside(i1["cell"])
        moved = int((before & ~after).sum())
        print(f"ions outside before {int(before.sum())}, re-injected {moved}", flush=True)
        assert moved > 16384
        assert sim.overflow() == 0


def test_batch_diagnostics_match_the_oracle():
    """Graph-window sums and timeData rows collected on the device against the
    oracle fed with the same per-step states: field and density averages,
    velocity slices and histograms bit-exact."""
    p = workloads.with_diagnostics(
        workloads.discharge(32, 40, mode=pic2d.MODE_PHYSICAL, mg_mode=pic2d.MG_FIXED, te=4.0),
        start=2, steps=4, time_data_freq=3, no_bins=301, bins_per_vt=30)
    N = p.nodes
    loc = list(p.dist_loc)
    sums = {k: np.zeros((N, N)) for k in ("rho", "phi", "ex", "ey", "q0", "q1")}
    d = [orc.Diag(N, p.no_bins), orc.Diag(N, p.no_bins)]
    rows = []
    qc = (p.electron_charge, p.ion_charge)
    with pic2d.Pic(p) as sim:
        for step in range(8):
            sim.advance(1)
            if step % 3 == 0:
                phi = sim.get_phi()
                qe3 = sim.get_q(0)[p.probe[5], p.probe[4]]
                qi3 = sim.get_q(1)[p.probe[5], p.probe[4]]
                rows.append([step, sim.particle_size(0), sim.particle_size(1),
                             phi[p.probe[1], p.probe[0]], phi[p.probe[3], p.probe[2]],
                             phi[p.probe[5], p.probe[4]], qe3, qi3])
            if 2 <= step < 6:
                ex, ey = sim.get_E()
                sums["rho"] += sim.get_rho()
                sums["phi"] += sim.get_phi()
                sums["ex"] += ex.reshape(-1)[: N * N].reshape(N, N)
                sums["ey"] += ey.reshape(-1)[: N * N].reshape(N, N)
                for s in (0, 1):
                    sums[f"q{s}"] += sim.get_q(s) / qc[s]
                    d[s].step(sim.get_state(s), p.vte if s == 0 else p.vti, p.divide_bins_by,
                              p.spread, loc)
        fa = sim.field_averages()
        for k in ("rho", "phi", "ex", "ey"):
            assert np.array_equal(fa[k], sums[k] / 4), k
        for s in (0, 1):
            assert np.array_equal(sim.density_average(s), sums[f"q{s}"] / 4)
            vx, vy = sim.velocity_slices(s)
            assert np.array_equal(vx, d[s].vpx / (p.vte * 4))
            assert np.array_equal(vy, d[s].vpy / (p.vte * 4))
            dist = sim.distributions(s, p.no_bins)
            assert np.array_equal(dist, d[s].dist.reshape(4, -1).astype(np.float64) / 4)
        assert d[0].dist.sum() > 0
        td = sim.time_data()
        assert np.array_equal(td, np.array(rows, dtype=np.float64))


def test_batch_output_files_have_the_reference_format(tmp_path):
    """pic_run writes the graph files at the end of the window and timeData.txt
    at the end of the run, with the line counts, field counts and number
    formats of the reference's example batch output (257x257 nodes)."""
    import json
    import os
    import re
    gold = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                       "batch_output_format.json")))
    p = workloads.with_diagnostics(
        workloads.discharge(256, 1, mode=pic2d.MODE_REF, mg_mode=pic2d.MG_FIXED),
        start=0, steps=2, time_data_freq=1)
    with pic2d.Pic(p) as sim:
        sim.set_output_dir(tmp_path)
        sim.run(3)
    efmt = re.compile(r"^-?\d\.\d{6}e[+-]\d{2}$")
    for name, g in gold.items():
        path = tmp_path / name
        assert path.exists(), name
        lines = path.read_text().split("\n")
        assert lines[-1] == ""
        lines = lines[:-1]
        if name == "timeData.txt":
            assert len(lines) == 3 // 1 + 1          # SIZE_TIME_DATA = run / freq + 1
        else:
            assert len(lines) == g["lines"], name
        for ln in lines[:3]:
            f = ln.split()
            assert len(f) == g["fields"], (name, ln)
            if name != "timeData.txt":
                assert all(efmt.match(x) for x in f), (name, ln)
                # trailing-space convention of each writer
                assert ln.endswith(" ") == g["trailing_space"], name
        if name == "timeData.txt":
            f0 = lines[0].split()   # recorded after step 0: step, then integer counts
            assert f0[0] == "0" and f0[1].isdigit() and f0[2].isdigit()
            assert int(f0[2]) <= p.init_particle_no


def test_snapshot_subset_and_graph_maxima(tmp_path):
    p = workloads.with_diagnostics(
        workloads.discharge(32, 20, mode=pic2d.MODE_PHYSICAL, mg_mode=pic2d.MG_FIXED),
        start=0, steps=2, time_data_freq=0, no_bins=101, bins_per_vt=10)
    with pic2d.Pic(p) as sim:
        sim.run(3)
        for s in (0, 1):
            st = sim.get_state(s)
            want = orc.state_to_aos({k: v[st["id"] % 7 == 0] for k, v in st.items()}, p.dx)
            got = sim.particle_subset(s, 7)
            key = lambda a: np.lexsort((a["vx"], a["y"], a["x"]))
            assert len(got) == len(want) > 0
            for f in ("x", "y", "vx", "vy", "vz", "weight"):
                assert np.array_equal(got[f][key(got)], want[f][key(want)]), f
        fa = sim.field_averages()
        mx = sim.graph_max()
        assert mx[0] == max(fa["phi"].max(), fa["phi"][0, 0])
        assert mx[1] == fa["ex"].max() and mx[2] == fa["ey"].max()
        assert mx[3] == sim.density_average(0).max()
        sim.write_snapshot(tmp_path, stride=5)
        step = sim.get_step()
        phi = np.loadtxt(tmp_path / f"snapshot_{step}_phi.txt")
        assert phi.shape == (p.nodes, p.nodes)
        np.testing.assert_allclose(phi, sim.get_phi(), rtol=1e-6, atol=1e-12)
        e = np.loadtxt(tmp_path / f"snapshot_{step}_e.txt")
        assert e.shape == (len(sim.particle_subset(0, 5)), 6)


def test_absorbing_compaction_keeps_store_order():
    """The absorbing walls' compaction (pic_boundary, pic_run's pre-loop
    pass) keeps the survivors in store order, so a sorted store stays
    sorted, and counts every removed particle at its wall.  Particles are
    moved outside on both x walls and the bottom wall in a pattern that
    spans many 1024-particle blocks."""
    p = workloads.discharge(64, 40, te=6.0)
    N = p.nodes
    print(f"absorbing compaction: {N}^2 nodes, {p.init_particle_no} particles", flush=True)
    with pic2d.Pic(p) as sim:
        d = sim.get_state(0)
        n = len(d["id"])
        k = np.arange(n)
        jx, jy = pic2d.cell_unpack(d["cell"])
        left, right = k % 7 == 0, (k % 11 == 0) & (k % 7 != 0)
        bottom = (k % 13 == 0) & ~left & ~right
        jx = np.where(left, -1, np.where(right, N - 1, jx))
        jy = np.where(bottom, -1, jy)
        d["cell"] = pic2d.cell_pack(jx, jy)
        sim.set_state(0, d)
        l0 = sim.loss_counts(0)
        sim.boundary(1)  # the electron wall
        out = sim.get_state(0)
        dl = sim.loss_counts(0) - l0
    keep = ~(left | right | bottom)
    assert np.array_equal(out["id"], d["id"][keep])
    for key in ("cell", "ox", "oy", "vx", "vy", "vz"):
        assert np.array_equal(out[key], d[key][keep]), key
    assert dl[0] == left.sum() and dl[1] == right.sum() and dl[2] == bottom.sum() and dl[3] == 0
