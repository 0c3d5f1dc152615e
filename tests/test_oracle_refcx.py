"""The reference-mode ion charge exchange of the oracle (orc32_ref_cx).

The per-ion probability is pinned by its closed form 1 - exp(-|v| dt 1e-20
n_g sigma) (sigma constant) and by the reference fit's sample values
sigma_cx(1 eV) = 77.0 and sigma_cx(10 eV) = 37.4 (1e-20 m^2); the accumulator
by a literal restatement of ion_collisions.cpp:297-303 run on the same
per-ion probabilities in (column, id) order, across calls (the accumulator
persists); the neutral table by its moments (vz >= 0 by construction)."""
import math

import numpy as np

import orc

Q, MI = 1.602177e-19, 39.948 * 1.66054e-27
N = 17


def params(sig_energy=1, sig_const=0.0, dt=1e-8, ng=3.2e21):
    return orc.RefCxParams(gas=0, sig_energy=sig_energy, sig_const=sig_const, q=Q, mass=MI, dt=dt,
                           ng=ng, key0=5, key1=6)


def test_probability_closed_form_and_fit_samples():
    c = params(sig_energy=0, sig_const=40.0)
    for v in (100.0, 1000.0, 5000.0):
        want = 1 - math.exp(-v * c.dt * 1e-20 * c.ng * 40.0)
        got = orc.lib().orc32_ref_cx_prob(orc.C.byref(c), v, 0.0, 0.0)
        # fp32: 1 - exp(-x) carries the rounding of exp near 1 (2^-24 absolute)
        assert abs(got - want) <= 2e-6 * want + 2.0 ** -23
    c = params()
    for E, sig in ((1.0, 77.0), (10.0, 37.4)):
        v = math.sqrt(2 * Q * E / MI)
        want = 1 - math.exp(-v * c.dt * 1e-20 * c.ng * sig)
        got = orc.lib().orc32_ref_cx_prob(orc.C.byref(c), 0.0, v, 0.0)
        assert abs(got - want) <= 2e-3 * want   # sample values are quoted to 3 digits


def test_accumulator_matches_the_reference_loop():
    rng = np.random.default_rng(3)
    n = 3000
    jx = rng.integers(-2, N + 2, n)
    jy = rng.integers(-2, N + 2, n)
    st = dict(cell=orc.cell_pack(jx, jy), ox=rng.random(n).astype(np.float32),
              oy=rng.random(n).astype(np.float32),
              vx=(rng.standard_normal(n) * 3000).astype(np.float32),
              vy=(rng.standard_normal(n) * 3000).astype(np.float32),
              vz=(rng.standard_normal(n) * 3000).astype(np.float32),
              id=rng.permutation(n).astype(np.uint32))
    c = params()
    table = orc.neutral_table(77, 250.0)
    pcol = np.zeros(N, np.float32)
    ref_p = np.zeros(N, np.float32)
    for step in range(3):
        out, ev = orc.ref_cx(st, N, c, pcol, table, step)
        # literal loop over ions in (column, id) order
        fired = set()
        order = []
        for k in range(n):
            x, y = st["ox"][k], st["oy"][k]
            inx = -1 <= jx[k] <= N - 1 and not (jx[k] == -1 and x == 0)
            iny = -1 <= jy[k] <= N - 1 and not (jy[k] == -1 and y == 0)
            if inx and iny:
                order.append((max(jx[k], 0), int(st["id"][k]), k))
        for col, _, k in sorted(order):
            P = orc.lib().orc32_ref_cx_prob(orc.C.byref(c), float(st["vx"][k]), float(st["vy"][k]),
                                            float(st["vz"][k]))
            ref_p[col] = np.float32(ref_p[col] + np.float32(P))
            if ref_p[col] > 1.0:
                fired.add(k)
                ref_p[col] = 0.0
        changed = {k for k in range(n) if out["vx"][k] != st["vx"][k]}
        assert changed == fired and ev == len(fired) and ev > 0
        assert np.array_equal(pcol, ref_p)
        st = out


def test_neutral_table_moments():
    vtn = 250.0
    t = orc.neutral_table(123, vtn).astype(np.float64).reshape(3, -1)
    n = t.shape[1]
    assert np.all(t[2] >= 0)                      # vz = vtn sin(pi R2) sqrt|ln R1| >= 0
    for comp in (t[0], t[1]):
        assert abs(comp.mean()) < 5 * vtn / math.sqrt(2 * n)
        assert abs(comp.var() / (vtn * vtn / 2) - 1) < 0.06
