"""Pins of the oracle's particle-count control (oracle/orc_control.c):

* the source (particle.cpp:491-510, velocities particle.cpp:452-464) against
  the distributions its formulas define: positions uniform on [dx, L-dx],
  v = vt cos(2 pi R2) sqrt|ln R1| is normal with variance vt^2/2;
* the reduction (particle.cpp:520-653) against a literal restatement of the
  reference's reduce_population loop (rand() % (size-1), swap and pop): group
  counts, quotas, survivors per group and their weights must agree exactly;
  the choice of removed particles is random in both and is checked for
  uniformity instead;
* the weighted deposit (pic.cl:46-49) against the unweighted one.
"""
import math
import random

import numpy as np
import pytest
from scipy import stats

import orc


L = 0.01
N = 65
DX = L / (N - 1)
VT = 5.0e5


def positions(d):
    jx, jy = orc.cell_unpack(d["cell"])
    return (jx + d["ox"].astype(np.float64)) * DX, (jy + d["oy"].astype(np.float64)) * DX


def test_inject_positions_uniform_in_inner_box():
    d = orc.inject(100000, 0, 7, 12345, DX, L, VT)
    x, y = positions(d)
    assert x.min() >= DX * (1 - 1e-6) and x.max() <= (L - DX) * (1 + 1e-6)
    assert y.min() >= DX * (1 - 1e-6) and y.max() <= (L - DX) * (1 + 1e-6)
    for c in (x, y):
        assert stats.kstest((c - DX) / (L - 2 * DX), "uniform").pvalue > 1e-3
    assert abs(np.corrcoef(x, y)[0, 1]) < 0.02
    assert np.all((d["ox"] >= 0) & (d["ox"] < 1) & (d["oy"] >= 0) & (d["oy"] < 1))


def test_inject_velocity_law():
    # vt cos(2 pi U2) sqrt(-ln U1) = (vt / sqrt 2) * (Box-Muller normal)
    d = orc.inject(100000, 1, 3, 999, DX, L, VT)
    for k in ("vx", "vy", "vz"):
        v = d[k].astype(np.float64) / (VT / math.sqrt(2.0))
        assert stats.kstest(v, "norm").pvalue > 1e-3, k
    m = np.vstack([d["vx"], d["vy"], d["vz"]]).astype(np.float64)
    c = np.corrcoef(m)
    assert np.all(np.abs(c[np.triu_indices(3, 1)]) < 0.02)


def test_inject_streams_differ_by_step_species_and_seed():
    a = orc.inject(64, 0, 1, 5, DX, L, VT)
    for other in (orc.inject(64, 0, 2, 5, DX, L, VT), orc.inject(64, 1, 1, 5, DX, L, VT),
                  orc.inject(64, 0, 1, 6, DX, L, VT)):
        assert not np.array_equal(a["vx"], other["vx"])
        assert not np.array_equal(a["ox"], other["ox"])
    b = orc.inject(64, 0, 1, 5, DX, L, VT)
    assert all(np.array_equal(a[k], b[k]) for k in a)


def params(ng=3, size_check=0, glimit=(300, 300, 300), gnew=(200, 200, 200), vt=VT):
    c = orc.ReduceParams(ng=ng, size_check=size_check, species=0, key0=77, key1=0)
    vl = (2 * vt, 3 * vt, 1000 * vt)
    for g in range(ng):
        c.vlim[g] = vl[g]
        c.glimit[g] = glimit[g]
        c.gnew[g] = gnew[g]
    return c


def ref_reduce_population(vx, vy, w, c, rng):
    """particle.cpp:538-632 spelled out: count, quotas, random removal by
    rand() % (size - 1) with swap-and-pop, group weights, re-weighting"""
    ng = c.ng
    vlim = [c.vlim[g] for g in range(ng)]
    parts = [[float(a), float(b), float(x)] for a, b, x in zip(vx, vy, w)]

    def group(p):
        v = math.sqrt(p[0] * p[0] + p[1] * p[1])
        g = 0
        while g < ng and v > vlim[g]:
            g += 1
        return g

    Ngroup = [0] * (ng + 1)
    for p in parts:
        Ngroup[group(p)] += 1
    Nremove = [0] * (ng + 1)
    for g in range(ng):
        if Ngroup[g] > c.glimit[g]:
            Nremove[g] = Ngroup[g] - c.gnew[g]
    total = sum(Nremove)
    j = [0] * (ng + 1)
    i = 0
    while i < total:
        R = rng.randrange(2 ** 31) % (len(parts) - 1)
        g = group(parts[R])
        if j[g] < Nremove[g]:
            j[g] += 1
            parts[R], parts[-1] = parts[-1], parts[R]
            parts.pop()
            i += 1
    gw = [1.0] * (ng + 1)
    if total > 0:
        for g in range(ng):
            if j[g] > 0:
                gw[g] = Ngroup[g] / (Ngroup[g] - j[g])
        for p in parts:
            p[2] = p[2] * gw[group(p)]
    return parts, Ngroup[:ng], Nremove[:ng]


def maxwell2d(n, seed):
    rng = np.random.default_rng(seed)
    v = (rng.standard_normal((2, n)) * VT).astype(np.float32)
    return v[0].copy(), v[1].copy()


@pytest.mark.parametrize("seed", [1, 2])
def test_reduce_matches_reference_loop(seed):
    n = 4000
    vx, vy = maxwell2d(n, seed)
    w = np.ones(n, dtype=np.float32)
    c = params(glimit=(1500, 900, 300), gnew=(1000, 600, 200))
    keep, w1, ng, nr, removed = orc.reduce(c, 5, vx, vy, np.arange(n), w)
    parts, Ngroup, Nremove = ref_reduce_population(vx, vy, w, c, random.Random(seed))
    assert list(ng[:3]) == Ngroup
    assert list(nr[:3]) == Nremove
    assert removed == sum(Nremove) > 0
    assert keep.sum() == len(parts)

    def grp(vx_, vy_):
        v = np.sqrt(vx_.astype(np.float64) ** 2 + vy_.astype(np.float64) ** 2)
        return np.searchsorted(np.array([c.vlim[g] for g in range(3)]), v, side="left")

    g_or = grp(vx[keep], vy[keep])
    ref = np.array(parts)
    g_ref = grp(ref[:, 0].astype(np.float32), ref[:, 1].astype(np.float32))
    for g in range(4):
        a = np.sort(w1[keep][g_or == g])
        b = np.sort(ref[:, 2][g_ref == g].astype(np.float32))
        assert np.array_equal(a, b), g
    # weight of each group is conserved (all weights 1 before)
    for g in range(3):
        if nr[g]:
            tot = w1[keep][g_or == g].astype(np.float64).sum()
            assert abs(tot - ng[g]) < 1e-4 * ng[g]


def test_reduce_mixed_weights_scale_per_particle():
    n = 3000
    vx, vy = maxwell2d(n, 3)
    rng = np.random.default_rng(3)
    w = rng.uniform(0.5, 2.0, n).astype(np.float32)
    c = params(glimit=(1000, 1000, 1000), gnew=(700, 700, 700))
    keep, w1, ng, nr, _ = orc.reduce(c, 1, vx, vy, np.arange(n), w)
    v = np.sqrt(vx.astype(np.float64) ** 2 + vy.astype(np.float64) ** 2)
    g = np.searchsorted(np.array([c.vlim[k] for k in range(3)]), v, side="left")
    for k in range(3):
        sel = keep & (g == k)
        gw = ng[k] / (ng[k] - nr[k]) if nr[k] else 1.0
        assert np.array_equal(w1[sel], (w[sel].astype(np.float64) * gw).astype(np.float32))
        assert (~keep & (g == k)).sum() == nr[k]


def test_reduce_gates_and_untouched_tail():
    n = 2000
    vx, vy = maxwell2d(n, 4)
    vx[:5] = 2000 * VT  # faster than the last limit: in no group
    w = np.ones(n, dtype=np.float32)
    # species not above size_to_check: nothing happens
    keep, w1, _, _, r = orc.reduce(params(size_check=n, glimit=(10, 10, 10), gnew=(5, 5, 5)), 1,
                                   vx, vy, np.arange(n), w)
    assert r == 0 and keep.all() and np.array_equal(w1, w)
    # no group above its limit: nothing happens
    keep, w1, ng, nr, r = orc.reduce(params(glimit=(n, n, n)), 1, vx, vy, np.arange(n), w)
    assert r == 0 and keep.all() and not nr.any() and ng[:3].sum() == n - 5
    # reduction: the tail keeps its particles and weights
    keep, w1, ng, nr, r = orc.reduce(params(glimit=(100, 100, 100), gnew=(50, 50, 50)), 1, vx,
                                     vy, np.arange(n), w)
    assert keep[:5].all() and np.all(w1[:5] == 1.0)
    assert r == nr.sum() > 0


def test_reduce_selection_uniform():
    # removal frequency of every particle over many independent steps
    n, trials = 1000, 400
    vx = np.full(n, 0.1 * VT, dtype=np.float32)
    vy = np.zeros(n, dtype=np.float32)
    c = params(ng=1, glimit=(500,), gnew=(750,))
    c.gnew[0] = 750
    hits = np.zeros(n)
    for t in range(trials):
        keep, *_ = orc.reduce(c, t, vx, vy, np.arange(n), np.ones(n, np.float32))
        assert (~keep).sum() == 250
        hits += ~keep
    # 10 id blocks: chi-square against equal removal rates
    blocks = hits.reshape(10, -1).sum(axis=1)
    assert stats.chisquare(blocks).pvalue > 1e-3
    assert abs(hits.mean() / trials - 0.25) < 0.01


def test_weighted_deposit():
    st = orc.random_state(5000, N, VT, 11)
    q1 = orc.deposit32(st, N, orc.MODE_PHYS)
    one = np.ones(5000, dtype=np.float32)
    assert np.array_equal(orc.deposit32_w(st, one, N, orc.MODE_PHYS), q1)
    assert np.array_equal(orc.deposit32_w(st, 2 * one, N, orc.MODE_PHYS), 2 * q1)
    assert not orc.deposit32_w(st, 0 * one, N, orc.MODE_PHYS).any()
    w = np.random.default_rng(2).uniform(0.1, 3.0, 5000).astype(np.float32)
    qw = orc.deposit32_w(st, w, N, orc.MODE_PHYS)
    # linear in the weight up to one rounding per corner
    acc = np.zeros(N * N)
    cnt = np.zeros(N * N)
    for k in range(5000):
        sub = {key: v[k:k + 1] for key, v in st.items()}
        qk = orc.deposit32(sub, N, orc.MODE_PHYS).ravel()
        acc += qk.astype(np.float64) * float(w[k])
        cnt += qk != 0
    assert np.all(np.abs(qw.ravel() - acc) <= 0.5 * cnt + 1e-6 * np.abs(acc))
