"""Exact resume: k steps, pic_checkpoint_write, a new handle filled by
pic_checkpoint_read, k more steps -- against 2k uninterrupted steps of the
same run: every particle (by id), counters, fixed-point
charge, phi, E, the solver's statistics, collision and loss counts, the
diagnostics' window sums and timeData rows, all bit for bit.  Covered: the
tolerance solve (its warm start reads the previous two potentials), argon
MCC with pair re-injection and energy-binned bounds, the reference's
steady walls, per-particle weights (source and reduction on), and a graph
window that spans the checkpoint."""
import os

import numpy as np
import pytest

import pic2d
import workloads

pytestmark = pytest.mark.gpu


def snapshot(sim):
    out = {}
    for s in (0, 1):
        st = sim.get_state(s)
        # the slot order within a sort bin follows the device's atomics, so
        # two runs agree particle by particle, not slot by slot
        o = np.argsort(st["id"], kind="stable")
        for k, v in st.items():
            out[f"{s}.{k}"] = v[o]
        out[f"{s}.q"] = sim.get_qfixed(s)
        out[f"{s}.lost"] = sim.loss_counts(s)
        out[f"{s}.coll"] = sim.collision_counts(s)
        out[f"{s}.n"] = np.array([sim.particle_count(s)])
    out["phi"] = sim.get_phi()
    out["E32"] = sim.get_E32()
    out["step"] = np.array([sim.get_step()])
    out["mg"] = np.array(sim.mg_stats(), dtype=np.float64)
    if sim.params.time_data_freq > 0:
        out["td"] = sim.time_data()
    if sim.params.graph_steps > 0:
        for k, v in sim.field_averages().items():
            out[f"avg.{k}"] = v
        out["dens0"] = sim.density_average(0)
    return out


def case(name):
    if name == "tol":
        p = workloads.discharge(64, 20, mode=pic2d.MODE_PHYSICAL, mg_mode=pic2d.MG_TOL, te=4.0)
        p.mg_tol = 1e-9
    elif name == "argon":
        p = workloads.with_argon(workloads.discharge(64, 20, te=4.0), pressure_pa=60.0)
    elif name == "ref_steady":
        p = workloads.discharge(32, 40, mode=pic2d.MODE_REF, mg_mode=pic2d.MG_FIXED, te=8.0)
        p.wall = pic2d.WALL_STEADY
        p.sigma_load_i = 3.0e5
    elif name == "weighted_diag":
        p = workloads.discharge(32, 30, mode=pic2d.MODE_PHYSICAL, mg_mode=pic2d.MG_TOL)
        p.number_to_inject = 40
        p.reduce = 1
        p.no_groups = 2
        for s in (0, 1):
            p.size_to_check[s] = 20000
            for g in range(2):
                p.v_limit[s][g] = (p.vte if s == 0 else p.vti) * (1.5 if g == 0 else 1e3)
                p.group_limit[s][g] = 12000
                p.new_group_size[s][g] = 9000
        workloads.with_diagnostics(p, start=5, steps=10, time_data_freq=2, no_bins=101,
                                   bins_per_vt=10)
        p.write_graphs = 0
    return p


@pytest.mark.parametrize("name", ["tol", "argon", "ref_steady", "weighted_diag"])
def test_resume_is_bit_exact(name, tmp_path):
    k = 8
    print(f"checkpoint {name}", flush=True)
    p = case(name)
    with pic2d.Pic(p) as a:
        a.run(2 * k)
        want = snapshot(a)
    path = tmp_path / "run.ckpt"
    with pic2d.Pic(p) as b:
        b.run(k)
        b.checkpoint(path)
    assert os.path.getsize(path) > 0
    with pic2d.Pic.resume(path) as c:
        assert c.get_step() == k
        c.advance(k)
        got = snapshot(c)
    assert want.keys() == got.keys()
    for key in want:
        assert np.array_equal(want[key], got[key]), (name, key)


def test_resume_rejects_other_parameters_and_corruption(tmp_path):
    p = case("tol")
    path = tmp_path / "run.ckpt"
    with pic2d.Pic(p) as a:
        a.run(3)
        a.checkpoint(path)
    q = pic2d.checkpoint_params(path)
    assert q.nodes == p.nodes and q.mg_tol == p.mg_tol and q.seed == p.seed
    other = case("tol")
    other.seed += 1
    with pytest.raises(pic2d.PicError) as e:
        pic2d.Pic.resume(path, params=other)
    assert e.value.code == pic2d.PIC_ERR_ARG
    # run_time may differ
    longer = case("tol")
    longer.run_time = 12345
    with pic2d.Pic.resume(path, params=longer) as c:
        assert c.get_step() == 3
    data = bytearray(path.read_bytes())
    data[len(data) // 2] ^= 0x40
    bad = tmp_path / "bad.ckpt"
    bad.write_bytes(bytes(data))
    with pytest.raises(pic2d.PicError) as e:
        pic2d.Pic.resume(bad)
    assert e.value.code == pic2d.PIC_ERR_FORMAT
    short = tmp_path / "short.ckpt"
    short.write_bytes(bytes(data[: len(data) // 3]))
    with pytest.raises(pic2d.PicError) as e:
        pic2d.Pic.resume(short)
    assert e.value.code == pic2d.PIC_ERR_FORMAT
