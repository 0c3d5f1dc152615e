"""Oracle particle functions pinned by closed forms (CIC weights, analytic
leapfrog, B rotation invariants, wall thresholds, charge conservation) and the
build-layout restatement (orc32_*) pinned against the FP64 reference one."""
import math

import numpy as np
import pytest

import orc

N, L = 33, 0.01
DX = L / (N - 1)


def aos(rows):
    a = np.zeros(len(rows), dtype=orc.PARTICLE_DTYPE)
    for k, r in enumerate(rows):
        a[k] = tuple(r) + (1.0,) * (6 - len(r))
    return a


def ref_gather(p, Ex, Ey):
    n = len(p)
    ex, ey = np.zeros(n), np.zeros(n)
    orc.lib().orc_ref_gather(orc.P(p), n, orc.P(Ex), orc.P(Ey), DX, DX, N, N, orc.P(ex), orc.P(ey))
    return ex, ey


def test_cic_single_particle_weights():
    x, y = (5 + 0.25) * DX, (7 + 0.6) * DX
    p = aos([(x, y, 0, 0, 0, 1.0)])
    q = np.zeros(N * N)
    orc.lib().orc_ref_deposit(orc.P(p), 1, 1.0, DX, DX, N, N, orc.P(q))
    q = q.reshape(N, N)
    wj, wk = x / DX - 5, y / DX - 7
    assert q[7, 5] == pytest.approx((1 - wj) * (1 - wk), rel=1e-12)   # a4 at (kp, jp)
    assert q[7, 6] == pytest.approx(wj * (1 - wk), rel=1e-12)         # a3 at (kp, jp+1)
    assert q[8, 5] == pytest.approx(wk * (1 - wj), rel=1e-12)         # a2 at (kp+1, jp)
    assert q[8, 6] == pytest.approx(wj * wk, rel=1e-12)               # a1
    assert q.sum() == pytest.approx(1.0, rel=1e-14)


def test_fixed_point_charge_is_exactly_conserved():
    st = orc.random_state(5000, N, 1.0, 3)
    q = orc.deposit32(st, N, orc.MODE_PHYS)
    one = 1 << orc.FX_TOTAL_BITS
    assert int(q.sum()) == 5000 * one
    # and agrees with the FP64 reference deposit to the 2^-24 weight
    # quantisation: per corner |d(wx wy)| <= (|dwx| + |dwy| + |dwx dwy|) < 2^-24
    # (each offset rounds by at most 2^-25), so a node is off by less than
    # 2^-24 times the particles that touch it
    p = orc.state_to_aos(st, DX)
    qr = np.zeros(N * N)
    orc.lib().orc_ref_deposit(orc.P(p), len(p), 1.0, DX, DX, N, N, orc.P(qr))
    jx, jy = orc.cell_unpack(st["cell"])
    touch = np.zeros((N, N))
    for dy in (0, 1):
        for dx in (0, 1):
            np.add.at(touch, (jy.astype(int) + dy, jx.astype(int) + dx), 1)
    err = np.abs(q.ravel() / float(one) - qr)
    assert np.all(err <= touch.ravel() * 2.0 ** -24 + 1e-15)
    assert err.max() > 0  # the quantisation is visible: the test is not vacuous


def test_uniform_field_gathers_uniformly():
    Ex = np.full((N + 1) * (N + 1), 3.0)
    Ey = np.full((N + 1) * (N + 1), -2.0)
    rng = np.random.default_rng(1)
    xy = (1 + rng.random((2, 100)) * (N - 3)) * DX
    p = aos([(a, b, 0, 0, 0) for a, b in xy.T])
    ex, ey = ref_gather(p, Ex, Ey)
    assert np.allclose(ex, 3.0, rtol=1e-14) and np.allclose(ey, -2.0, rtol=1e-14)
    # outside (x <= 0 or x >= N*dx): field stays zero (pic.cl:70-74)
    p = aos([(0.0, 5 * DX, 0, 0, 0), (-0.5 * DX, 5 * DX, 0, 0, 0), (N * DX + 1e-9, 5 * DX, 0, 0, 0)])
    ex, _ = ref_gather(p, Ex, Ey)
    assert np.all(ex == 0)


def test_constant_field_leapfrog_is_analytic():
    dt, coef, E = 1e-10, 5e3, 2.0
    x0, v0 = 10.3 * DX, 1e4
    p = aos([(x0, 10.5 * DX, v0, 0.0, 0.0)])
    exp, eyp = np.full(1, E), np.zeros(1)
    k = 7
    for _ in range(k):
        orc.lib().orc_ref_push(orc.P(p), 1, orc.P(exp), orc.P(eyp), DX, DX, N, N, coef, 1, 0, dt,
                               1, 0, 0.0)
    a = 2 * coef * E       # velocity increment per step
    assert p["vx"][0] == pytest.approx(v0 + k * a, rel=1e-12)
    assert p["x"][0] == pytest.approx(x0 + dt * (k * v0 + a * k * (k + 1) / 2), rel=1e-12)


def test_magnetic_rotation_preserves_speed_components():
    rng = np.random.default_rng(2)
    rows = [(15.5 * DX, 15.5 * DX, *rng.standard_normal(3) * 1e5) for _ in range(50)]
    p = aos(rows)
    ang, w = math.radians(30), 0.3
    ca, sa = math.cos(ang), math.sin(ang)
    vpar0 = ca * p["vx"] + sa * p["vy"]
    vperp0 = np.hypot(-sa * p["vx"] + ca * p["vy"], p["vz"])
    z = np.zeros(50)
    orc.lib().orc_ref_push(orc.P(p), 50, orc.P(z), orc.P(z), DX, DX, N, N, 0.0, ca, sa, 1e-14,
                           math.cos(w), math.sin(w), 1.0)
    assert np.allclose(ca * p["vx"] + sa * p["vy"], vpar0, rtol=1e-12)
    assert np.allclose(np.hypot(-sa * p["vx"] + ca * p["vy"], p["vz"]), vperp0, rtol=1e-12)


def test_absorbing_wall_thresholds():
    # lost iff (int)(x/dx) < 0 or >= N  (boundary.cpp:599): x <= -dx or x >= N*dx
    cases = [(-DX * 1.0000001, True), (-DX, True), (-0.5 * DX, False), (0.0, False),
             ((N - 1) * DX, False), (N * DX * 0.9999999, False), (N * DX, True)]
    p = aos([(x, 5 * DX, 0, 0, 0) for x, _ in cases])
    ids = np.arange(len(cases), dtype=np.uint32)
    n = orc.lib().orc_ref_absorb(orc.P(p), orc.P(ids), len(p), DX, DX, N, N)
    kept = set(ids[:n].tolist())
    for k, (_, lost) in enumerate(cases):
        assert (k not in kept) == lost, cases[k]
    # the cell/offset restatement classifies the same positions the same way
    st = {"cell": [], "ox": [], "oy": []}
    for x, _ in cases:
        t = x / DX
        j = math.floor(t)
        st["cell"].append(int(orc.cell_pack(j, 5)))
        st["ox"].append(t - j)
        st["oy"].append(0.5)
    n = len(cases)
    s = dict(cell=np.array(st["cell"], np.uint32), ox=np.array(st["ox"], np.float32),
             oy=np.array(st["oy"], np.float32), vx=np.zeros(n, np.float32),
             vy=np.zeros(n, np.float32), vz=np.zeros(n, np.float32), id=np.arange(n, dtype=np.uint32))
    pp = orc.PushParams(N=N, mode=orc.MODE_REF, cE=0, dtdx=0, dtdy=0, bfield=0, cos_a=1, sin_a=0,
                        cos_w=1, sin_w=0)
    _, status = orc.push32(s, np.zeros(2 * ((N + 1) * N + 1), np.float32), pp)
    for k, (_, lost) in enumerate(cases):
        assert (status[k] != orc.ST_KEPT) == lost, cases[k]



@pytest.mark.parametrize("steps", [1, 10])
def test_fp32_cell_layout_tracks_fp64_reference(steps):
    """orc32_run (build layout) vs orc_ref_run (reference layout), same start."""
    import types
    n = 20000
    te, ti = 2.0, 0.026
    q, me, amu = 1.602177e-19, 9.10939e-31, 1.66054e-27
    mi = 39.948 * amu
    wpe = math.sqrt(1e15 * 2.566971e-38 / (8.854e-12 * me))
    p = types.SimpleNamespace(nodes=N, length=L, mode=orc.MODE_REF, corner8=1, mg_ncycle=2,
                              mg_max_iter=30, dt_dti=1, epsilon=8.854e-12, mg_tol=1e-8,
                              phi_left=0.0, phi_right=0.0, phi_bottom=0.0, phi_top=0.0,
                              phi_period=0.0, qce=-q * 1e15 * L * L / n, qci=q * 1e15 * L * L / n,
                              electron_charge=-q, electron_mass=me, ion_charge=q, ion_mass=mi,
                              dt=0.1 / wpe, dti=0.1 / wpe, magnetic_field=0.0, angle=0.5)
    c32, cref = orc.configs_from_params(p)
    e = orc.random_state(n, N, math.sqrt(q * te / me), 11)
    i = orc.random_state(n, N, math.sqrt(q * ti / mi), 12)
    E32 = np.zeros(2 * ((N + 1) * N + 1), np.float32)
    r32 = orc.run32(c32, steps, e, i, E32)
    Ex = np.zeros((N + 1) * (N + 1))
    rr = orc.run_ref(cref, steps, orc.state_to_aos(e, DX), orc.state_to_aos(i, DX), e["id"], i["id"],
                     Ex, Ex.copy())
    # positions and velocities of particles present in both, matched by id
    a = orc.state_to_aos(r32["e"], DX)
    ia = {int(k): j for j, k in enumerate(r32["e"]["id"])}
    sel = [(ia[int(k)], j) for j, k in enumerate(rr["eid"]) if int(k) in ia]
    assert len(sel) >= 0.99 * len(rr["eid"])
    ga = a[[s for s, _ in sel]]
    gb = rr["e"][[s for _, s in sel]]
    for f in ("x", "y", "vx", "vy"):
        rel = np.linalg.norm(ga[f] - gb[f]) / np.linalg.norm(gb[f])
        assert rel < 1e-6, (f, rel)
    rel_phi = np.linalg.norm(r32["phi"] - rr["phi"]) / np.linalg.norm(rr["phi"])
    assert rel_phi < 2e-5, rel_phi
    assert abs(len(r32["e"]["id"]) - len(rr["eid"])) <= 2


def test_dump_roundtrip_and_reference_format(tmp_path):
    rng = np.random.default_rng(4)
    p = np.zeros(100, dtype=orc.PARTICLE_DTYPE)
    for f in ("x", "y", "vx", "vy", "vz"):
        p[f] = rng.standard_normal(100) * 1e-3
    p["weight"] = 1.0
    f17 = str(tmp_path / "d17.txt")
    orc.lib().orc_dump_write(f17.encode(), orc.P(p), 100, 17)
    back = np.zeros(100, dtype=orc.PARTICLE_DTYPE)
    assert orc.lib().orc_dump_read(f17.encode(), orc.P(back), 100) == 100
    assert np.array_equal(back, p)
    f6 = str(tmp_path / "d6.txt")
    orc.lib().orc_dump_write(f6.encode(), orc.P(p), 100, 6)
    first = open(f6).readline().split()
    assert first == ["%g" % v for v in p[0].tolist()]       # ostream default precision
