"""Branches of the timestep that the default workloads never take, each
bit-exact against the build-layout oracle over 10 fused steps (cells,
offsets, velocities by id, fixed-point charge, phi, E, loss counts):

* the magnetic rotation of the push (pic.cl:121-137), B > 0 at an angle;
* ion sub-cycling, DT_DTI = 2 and 4 (pic.cpp:1097-1128: ions move on every
  DT_DTI-th step with DTI = DT_DTI * DT);
* Dirichlet walls at nonzero potentials with the right wall driven as
  PHI_RIGHT cos(PHI_PERIOD * step) (fields.cpp:177, 193-219);
* the reference's lattice loader (particle.cpp:410-443), the default of
  every reference parameters.h: positions against a restatement of the
  reference's loop, both species on the same points (zero initial charge).
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


def run_bit_exact(p, steps=10, label=""):
    """fused GPU steps against orc32_run from the GPU's initial state"""
    print(f"branch parity {label}: N={p.nodes} n={p.init_particle_no} steps={steps}", flush=True)
    with pic2d.Pic(p) as sim:
        e0, i0, E0 = sim.get_state(0), sim.get_state(1), sim.get_E32()
        c32, _ = orc.configs_from_params(p)
        c32.preloop = 1
        r = orc.run32(c32, steps, e0, i0, E0)
        sim.run(steps)
        assert_states_equal(sim.get_state(0), r["e"], f"{label} e")
        assert_states_equal(sim.get_state(1), r["i"], f"{label} i")
        assert np.array_equal(sim.get_qfixed(0), r["qe"])
        assert np.array_equal(sim.get_qfixed(1), r["qi"])
        assert np.array_equal(sim.get_phi(), r["phi"])
        assert np.array_equal(sim.get_E32(), r["E32"])
        lost = np.concatenate([sim.loss_counts(0)[:4], sim.loss_counts(1)[:4]])
        assert np.array_equal(lost, r["lost"])
        return r, (c32, e0, i0, E0)


@pytest.mark.parametrize("mode", [pic2d.MODE_REF, pic2d.MODE_PHYSICAL])
def test_magnetic_rotation(mode):
    p = workloads.discharge(64, 20, mode=mode, mg_mode=pic2d.MG_FIXED)
    # strong enough that the rotation angle per step is visible in fp32
    p.magnetic_field = 0.05
    p.angle = 0.3
    r, (c32, e0, i0, E0) = run_bit_exact(p, label=f"B mode={mode}")
    # the rotation is exercised: without it the same steps end elsewhere
    c32.pe.bfield = c32.pi.bfield = 0
    r0 = orc.run32(c32, 10, e0, i0, E0)
    a, b = by_id(r["e"]), by_id(r0["e"])
    common = np.intersect1d(a["id"], b["id"])
    assert np.mean(a["vz"][np.searchsorted(a["id"], common)] !=
                   b["vz"][np.searchsorted(b["id"], common)]) > 0.9


@pytest.mark.parametrize("k", [2, 4])
def test_ion_subcycling(k):
    p = workloads.discharge(64, 20, mode=pic2d.MODE_REF, mg_mode=pic2d.MG_FIXED, ti=2.0)
    p.dt_dti = k
    p.dti = k * p.dt
    run_bit_exact(p, label=f"DT_DTI={k}")


@pytest.mark.parametrize("mode", [pic2d.MODE_REF, pic2d.MODE_PHYSICAL])
def test_dirichlet_walls_and_driven_right_wall(mode):
    p = workloads.discharge(64, 20, mode=mode, mg_mode=pic2d.MG_FIXED)
    p.phi_left, p.phi_right, p.phi_bottom, p.phi_top = 1.5, -2.0, 0.5, 3.0
    p.phi_period = 0.3
    r, _ = run_bit_exact(p, label=f"phi BCs mode={mode}")
    phi = r["phi"]
    # walls hold the prescribed values (rows win at the corners); the right
    # wall is PHI_RIGHT cos(PHI_PERIOD * time_step) of the last step,
    # time_step = 9 (pic.cpp:1069, 1158)
    assert np.all(phi[1:-1, 0] == 1.5)
    assert np.allclose(phi[1:-1, -1], -2.0 * np.cos(0.3 * 9), rtol=0, atol=1e-15)
    assert np.all(phi[0, :] == 0.5) and np.all(phi[-1, :] == 3.0)


def reference_lattice(n, L):
    """particle.cpp:410-443 restated: rows of points dxp apart from dxp,
    dxp = XLENGTH / (sqrt(INIT_PARTICLE_NO) + 1), positions accumulated by
    repeated addition; a second layer starts over when the lattice is full"""
    dxp = L / (np.sqrt(float(n)) + 1.0)
    xs, ys = [], []
    i = 0
    while i < n:
        y = dxp
        while y < L and i < n:
            x = dxp
            while x < L and i < n:
                xs.append(x)
                ys.append(y)
                x += dxp
                i += 1
            y += dxp
    return np.array(xs), np.array(ys)


def test_lattice_loader():
    p = workloads.discharge(32, 16, mode=pic2d.MODE_REF, mg_mode=pic2d.MG_FIXED)
    p.loader = pic2d.LOAD_LATTICE
    n = p.init_particle_no
    dx = p.length / (p.nodes - 1)
    xs, ys = reference_lattice(n, p.length)
    with pic2d.Pic(p) as sim:
        e, i = by_id(sim.get_state(0)), by_id(sim.get_state(1))
        assert np.array_equal(e["id"], np.arange(n)) and np.array_equal(i["id"], np.arange(n))
        for s in (e, i):
            a = orc.state_to_aos(s, dx)
            # fp32 in-cell offsets: within half an ulp of the offset, times dx
            assert np.max(np.abs(a["x"] - xs)) <= 2.0 ** -24 * dx * 1.01
            assert np.max(np.abs(a["y"] - ys)) <= 2.0 ** -24 * dx * 1.01
        # both species on the same points: the first deposit cancels exactly
        for k in ("cell", "ox", "oy"):
            assert np.array_equal(e[k], i[k])
    run_bit_exact(p, label="lattice")
