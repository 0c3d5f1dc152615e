"""Pins of the oracle's surface physics (oracle/orc_wall.c) against what the
reference's formulas fix (boundary.cpp:261-529, 586-824):

* Vaughan's yield equals max_yield * (1 + k theta^2 / 2 pi) where w = 1 and 0
  at E0; the Ar+ yield is 0 below 80 eV and 0.006 E / (1 + E/10) at 80 eV;
* branch frequencies: an event with probability min(1, yield), then
  reflection and backscatter 0.1 yield / yield each above 50 eV, else
  floor(|0.8 yield|) + 1 true secondaries;
* emitted electrons: energy uniform on [0, thr - 2 wf] (Krimke eq. 7),
  sin^2 chi uniform (eq. 5), placed R dt |v_n| inside the wall;
* reference-mode quirks: angle from 2E/m with E in eV (about pi/2),
  secondaries at y = 0 (x walls) / x = 0 (y walls), wall at N dx.
"""
import math

import numpy as np
import pytest
from scipy import stats

import orc

N = 33
L = 0.01
DX = L / (N - 1)
DT = 5e-11
ME, MI = 9.10939e-31, 39.948 * 1.66054e-27
Q = orc.CHARGE


def params(mode=orc.MODE_PHYS, see_e=1, refl=1, back=1, see_i=1, seed=99):
    return orc.WallParams(mode=mode, N=N, dx=DX, Lw=L if mode == orc.MODE_PHYS else DX * N, dt=DT,
                          me=ME, mi=MI, q=Q, max_yield=1.4, E0=4.32, Emax0=650.0, thr=15.75,
                          wf=4.6, see_e=see_e, refl=refl, back=back, see_i=see_i, key0=seed,
                          key1=0)


def beam(n, E_eV, wall="left", mass=ME, mode=orc.MODE_PHYS, theta=0.0):
    """n particles just outside `wall`, moving into it at angle theta to the normal"""
    v = math.sqrt(2.0 * E_eV * Q / mass)
    vn, vt = v * math.cos(theta), v * math.sin(theta)
    rng = np.random.default_rng(1)
    along = rng.uniform(2, N - 3, n)
    jx = np.zeros(n, np.int32)
    jy = np.zeros(n, np.int32)
    ox = np.zeros(n, np.float32)
    oy = np.zeros(n, np.float32)
    vx = np.zeros(n, np.float32)
    vy = np.zeros(n, np.float32)
    out = -1 if mode == orc.MODE_PHYS else -2       # reference mode: x <= -dx
    far = N - 1 if mode == orc.MODE_PHYS else N     # reference mode: x >= N dx
    a_j = np.floor(along).astype(np.int32)
    a_o = (along - a_j).astype(np.float32)
    if wall in ("left", "right"):
        jx[:] = out if wall == "left" else far
        ox[:] = 0.5
        jy, oy = a_j, a_o
        vx[:] = -vn if wall == "left" else vn
        vy[:] = vt
    else:
        jy[:] = out if wall == "bottom" else far
        oy[:] = 0.5
        jx, ox = a_j, a_o
        vy[:] = -vn if wall == "bottom" else vn
        vx[:] = vt
    return dict(cell=orc.cell_pack(jx, jy), ox=ox, oy=oy, vx=vx, vy=vy,
                vz=np.zeros(n, np.float32), id=np.arange(n, dtype=np.uint32))


def test_vaughan_closed_forms():
    f = orc.lib().orc_vaughan
    assert f(650.0, 0.0, 1.4, 4.32, 650.0) == 1.4           # w = 1 at normal incidence
    assert f(4.32, 0.0, 1.4, 4.32, 650.0) == 0.0            # w = 0 at the threshold
    fac = 1.0 + 0.25 * (math.pi / 2) ** 2 / (2 * math.pi)
    assert abs(f(650.0 * fac, math.pi / 2, 1.4, 4.32, 650.0) - 1.4 * fac) < 1e-12
    E = np.linspace(5, 649, 50)
    y = [f(e, 0.0, 1.4, 4.32, 650.0) for e in E]
    assert np.all(np.diff(y) > 0)                            # rising up to Emax0
    assert math.isnan(f(2.0, 0.0, 1.4, 4.32, 650.0))        # below E0, as written


def test_ar_ion_yield_points():
    g = orc.lib().orc_ar_ion_yield
    assert g(79.999) == 0.0 and g(10.0) == 0.0
    assert abs(g(80.0) - 0.006 * 80 / 9.0) < 1e-15
    assert g(1000.0) > g(200.0) > g(80.0)


@pytest.mark.parametrize("wall", ["left", "right", "bottom", "top"])
def test_branch_frequencies(wall):
    n, E = 40000, 200.0
    st = beam(n, E, wall)
    c = params()
    out, cnt = orc.wall(c, 0, 3, st)
    y = orc.lib().orc_vaughan(E, 0.0, 1.4, 4.32, 650.0)
    assert cnt[0] == n
    p_ev = min(1.0, y)
    sd = math.sqrt(n * p_ev * (1 - p_ev)) + 1
    assert abs(cnt[2] + cnt[3] + cnt[4] - n * p_ev) < 5 * sd
    ev = cnt[2] + cnt[3] + cnt[4]
    for k, p in ((2, 0.1 * y / (y + 1e-6)), (3, 0.1 * y / (y + 1e-6))):
        assert abs(cnt[k] - ev * p) < 5 * math.sqrt(ev * p * (1 - p)) + 1
    per = int(math.floor(abs(0.8 * y))) + 1
    assert cnt[5] == per * cnt[4]
    assert len(out["key"]) == cnt[2] + cnt[3] + cnt[5]
    src = out["key"] >> np.uint64(32)
    assert (src == 4).sum() == cnt[2] and (src == 5).sum() == cnt[3]


def test_secondary_energy_and_angles():
    n = 30000
    c = params(refl=0, back=0)
    out, cnt = orc.wall(c, 0, 1, beam(n, 650.0, "left"))  # yield 1.4: every one emits
    assert cnt[4] == n and cnt[5] == 2 * n
    v2 = out["vx"].astype(np.float64) ** 2 + out["vy"] ** 2 + out["vz"] ** 2
    e = 0.5 * ME * v2 / Q / (15.75 - 2 * 4.6)
    assert stats.kstest(e, "uniform").pvalue > 1e-3
    assert np.all(out["vx"] >= 0)                           # into the plasma from x = 0
    sin2 = (out["vy"].astype(np.float64) ** 2 + out["vz"] ** 2) / v2
    assert stats.kstest(sin2, "uniform").pvalue > 1e-3       # chi = asin(sqrt(R))
    jx, _ = orc.cell_unpack(out["cell"])
    x = (jx + out["ox"].astype(np.float64)) * DX
    r = x / (DT * out["vx"].astype(np.float64))
    assert np.all((r >= 0) & (r <= 1 + 1e-5))
    assert stats.kstest(r, "uniform").pvalue > 1e-3


def test_reflection_reverses_normal_velocity():
    n = 20000
    c = params(see_e=1, refl=1, back=0)
    st = beam(n, 300.0, "top", theta=0.4)
    out, cnt = orc.wall(c, 0, 2, st)
    sel = (out["key"] >> np.uint64(32)) == 4
    ids = (out["key"][sel] & np.uint64(0xFFFFFFFF)).astype(np.int64)
    assert cnt[2] == sel.sum() > 0
    assert np.array_equal(out["vy"][sel], -st["vy"][ids])
    assert np.array_equal(out["vx"][sel], st["vx"][ids])
    _, jy = orc.cell_unpack(out["cell"][sel])
    y = (jy + out["oy"][sel].astype(np.float64)) * DX
    assert np.all((y <= L) & (y >= L - DT * np.abs(st["vy"][ids]) * (1 + 1e-6)))


def test_backscatter_speed_fraction():
    n = 40000
    c = params(see_e=1, refl=0, back=1)
    E = 300.0
    out, cnt = orc.wall(c, 0, 2, beam(n, E, "right"))
    sel = (out["key"] >> np.uint64(32)) == 5
    assert cnt[3] == sel.sum() > 500
    v = np.sqrt(out["vx"][sel].astype(np.float64) ** 2 + out["vy"][sel] ** 2 + out["vz"][sel] ** 2)
    frac = v / math.sqrt(2 * E * Q / ME)
    assert np.all(frac <= 1 + 1e-6)
    assert stats.kstest(frac, "uniform").pvalue > 1e-3      # (rand() % 10000) / 10000
    assert np.all(out["vx"][sel] <= 0)                      # back into the plasma from x = L


def test_reference_mode_quirks():
    n = 20000
    c = params(mode=orc.MODE_REF, refl=0, back=0)
    out, cnt = orc.wall(c, 0, 5, beam(n, 100.0, "left", mode=orc.MODE_REF))
    sec = (out["key"] >> np.uint64(32)) >= 8
    jx, jy = orc.cell_unpack(out["cell"][sec])
    assert np.all((jy == 0) & (out["oy"][sec] == 0))         # y left at 0.0
    # angle acos(vx / (sqrt(2 E / m) + 1e-6)) with E in eV: ~ pi/2
    th = math.acos(-math.sqrt(2 * 100.0 * Q / ME) / (math.sqrt(2 * 100.0 / ME) + 1e-6))
    y_ref = orc.lib().orc_vaughan(100.0, th, 1.4, 4.32, 650.0)
    assert abs(th - math.pi / 2) < 1e-8
    p = min(1.0, y_ref)
    assert abs(cnt[4] - n * p) < 5 * math.sqrt(n * p * (1 - p)) + 1
    # y walls: x left at 0.0; right/top walls re-emit from N dx
    out, _ = orc.wall(c, 0, 5, beam(n, 100.0, "top", mode=orc.MODE_REF))
    jx, jy = orc.cell_unpack(out["cell"])
    assert np.all((jx == 0) & (out["ox"] == 0))
    y = (jy + out["oy"].astype(np.float64)) * DX
    assert np.all((y <= DX * N) & (y > DX * (N - 1)))
    # below E0 the reference's yield is NaN: absorbed
    _, cnt = orc.wall(c, 0, 5, beam(1000, 2.0, "left", mode=orc.MODE_REF))
    assert cnt[1] == 1000


def test_ion_impact_emission():
    n = 40000
    c = params()
    E = 1000.0
    out, cnt = orc.wall(c, 1, 4, beam(n, E, "bottom", mass=MI))
    y = orc.lib().orc_ar_ion_yield(E)
    assert cnt[1] + cnt[4] == n and cnt[5] == cnt[4]
    assert abs(cnt[4] - n * y) < 5 * math.sqrt(n * y * (1 - y))
    assert np.all((out["key"] >> np.uint64(32)) == 16)
    v2 = out["vx"].astype(np.float64) ** 2 + out["vy"] ** 2 + out["vz"] ** 2
    e = 0.5 * ME * v2 / Q / (15.75 - 2 * 4.6)
    assert stats.kstest(e, "uniform").pvalue > 1e-3
    assert np.all(out["vy"] >= 0)
    # no emission below 80 eV or with the option off
    _, cnt = orc.wall(c, 1, 4, beam(1000, 50.0, "left", mass=MI))
    assert cnt[1] == 1000
    _, cnt = orc.wall(params(see_i=0), 1, 4, beam(1000, 5000.0, "left", mass=MI))
    assert cnt[1] == 1000


def test_options_off_absorb_everything():
    st = beam(5000, 300.0, "left")
    for c in (params(see_e=0), params(see_e=0, refl=1, back=1)):
        out, cnt = orc.wall(c, 0, 1, st)
        assert cnt[1] == 5000 and len(out["key"]) == 0


def test_draws_are_keyed():
    st = beam(2000, 300.0, "left")
    a, _ = orc.wall(params(), 0, 1, st)
    b, _ = orc.wall(params(), 0, 1, st)
    c, _ = orc.wall(params(), 0, 2, st)
    assert all(np.array_equal(a[k], b[k]) for k in a)
    assert not np.array_equal(a["key"], c["key"]) or not np.array_equal(a["vx"], c["vx"])
