"""Surface physics of the non-steady walls on the GPU against the oracle
(orc_wall.c; boundary.cpp:261-529, 586-824).

* Phase entry point (pic_boundary): electrons or ions placed outside every
  wall over 5-1000 eV (ions 50-3000 eV) are absorbed, reflected,
  backscattered or emit secondaries exactly as orc_wall decides; returned
  electrons keep their id, emitted ones take new ids in staging-key order.
  Cells and the branch taken are compared exactly; velocities and offsets
  pass through double sin/cos/asin/pow on both sides and are compared to
  2e-7 relative.
* Fused step: a hot-electron run's first step equals the oracle's push,
  then orc_wall on the electrons that left, then the commit; its charge is
  the fixed-point deposit of that population.
"""
import math

import numpy as np
import pytest

import orc
import pic2d
import workloads

pytestmark = pytest.mark.gpu

FKEYS = ("ox", "oy", "vx", "vy", "vz")


def by_id(s):
    o = np.argsort(s["id"], kind="stable")
    return {k: v[o] for k, v in s.items()}


def close(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.all(np.abs(a - b) <= 2e-7 * np.maximum(np.abs(a), np.abs(b)) + 1e-30)


def make(mode, te=2.0, ti=0.026, **kw):
    p = workloads.discharge(32, 10, mode=mode, mg_mode=pic2d.MG_FIXED, te=te, ti=ti)
    p.e_impact_see = 1
    p.ion_impact_see = 1
    p.reflection = 1
    p.backscatter = 1
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def outside_particles(n, N, mode, mass, emin, emax, seed, id0):
    """n particles outside one of the four walls (reference mode: x <= -dx or
    x >= N dx), moving outward with energies uniform in [emin, emax] eV"""
    rng = np.random.default_rng(seed)
    wall = rng.integers(0, 4, n)
    out_lo = -1 if mode == pic2d.MODE_PHYSICAL else -2
    out_hi = N - 1 if mode == pic2d.MODE_PHYSICAL else N
    along = rng.uniform(1, N - 2, n)
    aj = np.floor(along).astype(np.int32)
    ao = (along - aj).astype(np.float32)
    jx = np.where(wall == 0, out_lo, np.where(wall == 1, out_hi, aj)).astype(np.int32)
    jy = np.where(wall == 2, out_lo, np.where(wall == 3, out_hi, aj)).astype(np.int32)
    ox = np.where(wall < 2, np.float32(0.25), ao).astype(np.float32)
    oy = np.where(wall >= 2, np.float32(0.25), ao).astype(np.float32)
    E = rng.uniform(emin, emax, n)
    v = np.sqrt(2 * E * orc.CHARGE / mass)
    th = rng.uniform(0, 1.4, n)
    ph = rng.uniform(0, 2 * math.pi, n)
    vn, vt = v * np.cos(th), v * np.sin(th)
    sign = np.where((wall == 0) | (wall == 2), -1.0, 1.0)
    vx = np.where(wall < 2, sign * vn, vt * np.cos(ph))
    vy = np.where(wall >= 2, sign * vn, vt * np.cos(ph))
    vz = vt * np.sin(ph)
    return dict(cell=orc.cell_pack(jx, jy), ox=ox, oy=oy, vx=vx.astype(np.float32),
                vy=vy.astype(np.float32), vz=vz.astype(np.float32),
                id=(id0 + np.arange(n)).astype(np.uint32))


def expected_ids(keys, next_id):
    """ids the commit gives the staged particles: sources 4, 5 keep theirs,
    the others take next_id + rank in key order"""
    order = np.argsort(keys, kind="stable")
    ids = np.empty(len(keys), dtype=np.uint32)
    for r, k in enumerate(order):
        src = int(keys[k]) >> 32
        ids[k] = (int(keys[k]) & 0xFFFFFFFF) if src in (4, 5) else next_id + r
    return ids


def cat(a, b):
    return {k: np.concatenate([a[k], b[k]]) for k in a}


def check_set(got, want, what):
    got, want = by_id(got), by_id(want)
    assert np.array_equal(got["id"], want["id"]), what
    assert np.array_equal(got["cell"], want["cell"]), what
    for k in FKEYS:
        assert close(got[k], want[k]), (what, k)


@pytest.mark.parametrize("mode", [pic2d.MODE_REF, pic2d.MODE_PHYSICAL])
def test_electron_wall_phase_matches_oracle(mode):
    p = make(mode)
    print(f"electron wall phase: mode={mode}", flush=True)
    with pic2d.Pic(p) as sim:
        n0 = sim.particle_size(0)
        inside = {k: v[:3000] for k, v in sim.get_state(0).items()}
        inside["id"] = np.arange(3000, dtype=np.uint32)
        outs = outside_particles(6000, p.nodes, mode, p.electron_mass, 1.0, 1000.0, 3, 3000)
        st = cat(inside, outs)
        sim.set_state(0, st)
        sim.boundary(1)
        c = orc.wall_params(p)
        res, cnt = orc.wall(c, 0, 0, outs)
        assert cnt[2] > 50 and cnt[3] > 50 and cnt[4] > 500 and cnt[1] > 100, cnt
        next_id = max(n0, 9000)
        res["id"] = expected_ids(res["key"], next_id)
        want = cat(inside, {k: res[k] for k in ("cell", "ox", "oy", "vx", "vy", "vz", "id")})
        check_set(sim.get_state(0), want, f"electrons mode={mode}")
        assert np.array_equal(sim.wall_counts(0), cnt)
        assert sim.loss_counts(0)[:4].sum() == 6000


@pytest.mark.parametrize("mode", [pic2d.MODE_REF, pic2d.MODE_PHYSICAL])
def test_ion_wall_phase_matches_oracle(mode):
    p = make(mode)
    print(f"ion wall phase: mode={mode}", flush=True)
    with pic2d.Pic(p) as sim:
        ne0 = sim.particle_size(0)
        e0 = sim.get_state(0)
        ions_in = {k: v[:2000] for k, v in sim.get_state(1).items()}
        ions_in["id"] = np.arange(2000, dtype=np.uint32)
        outs = outside_particles(20000, p.nodes, mode, p.ion_mass, 50.0, 3000.0, 4, 2000)
        sim.set_state(1, cat(ions_in, outs))
        sim.boundary(0)
        c = orc.wall_params(p)
        res, cnt = orc.wall(c, 1, 0, outs)
        assert cnt[4] > 200, cnt
        res["id"] = expected_ids(res["key"], ne0)
        want_e = cat(e0, {k: res[k] for k in ("cell", "ox", "oy", "vx", "vy", "vz", "id")})
        check_set(sim.get_state(0), want_e, f"electrons after ion wall mode={mode}")
        check_set(sim.get_state(1), ions_in, "ions kept")
        assert np.array_equal(sim.wall_counts(1), cnt)


@pytest.mark.parametrize("mode", [pic2d.MODE_REF, pic2d.MODE_PHYSICAL])
@pytest.mark.parametrize("sort_interval", [1, 4])
def test_fused_step_with_surface_physics(mode, sort_interval):
    """sort_interval 1: the sort pass queues the wall particles; 4: the
    in-place pass does"""
    p = make(mode, te=1000.0, sort_interval=sort_interval)
    print(f"fused step with walls: mode={mode} sort_interval={sort_interval}", flush=True)
    c32, _ = orc.configs_from_params(p)
    with pic2d.Pic(p) as sim:
        e0 = sim.get_state(0)
        n0 = len(e0["id"])
        E0 = sim.get_E32()
        sim.run(1)
        moved, status = orc.push32(e0, E0, c32.pe)
        kept = {k: v[status == 0] for k, v in moved.items()}
        lost = {k: v[status != 0] for k, v in moved.items()}
        res, cnt = orc.wall(orc.wall_params(p), 0, 0, lost)
        assert cnt[0] > 100 and cnt[2] + cnt[3] + cnt[4] > 50, cnt
        res["id"] = expected_ids(res["key"], n0)
        ret = {k: res[k] for k in ("cell", "ox", "oy", "vx", "vy", "vz", "id")}
        got = sim.get_state(0)
        check_set(got, cat(kept, ret), f"electrons@1 mode={mode}")
        w = np.ones(len(got["id"]), np.float32)
        assert np.array_equal(sim.get_qfixed(0), orc.deposit32_w(got, w, p.nodes, c32.mode))
        e_counts = sim.wall_counts(0)
        assert np.array_equal(e_counts, cnt)
        # bookkeeping over more steps
        sim.run(4)
        ce, ci = sim.wall_counts(0), sim.wall_counts(1)
        assert ce[0] == ce[1] + ce[2] + ce[3] + ce[4]
        assert sim.loss_counts(0)[:4].sum() == ce[0]
        assert sim.loss_counts(1)[:4].sum() == ci[0]
        assert sim.particle_size(0) == n0 - ce[0] + ce[2] + ce[3] + ce[5] + ci[5]
