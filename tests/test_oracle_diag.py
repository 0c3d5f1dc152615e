"""The oracle's batch diagnostics (orc32_diag_particles), pinned against a
literal restatement of the reference kernels sum_av_velocities and
sum_av_dist (pic.cl:151-258) on absolute FP64 positions x = (jx + ox) dx with
jp = (int)(x/dx): the same cells and histogram bins are counted, and the
fixed-point velocity sums agree with the plain sums to the fixed-point
rounding (2^-21 m/s per particle)."""
import math

import numpy as np

import orc

N = 33
DX = 0.01 / (N - 1)


def reference(st, vt, nb, divide, spread, loc):
    x = (orc.cell_unpack(st["cell"])[0] + st["ox"].astype(np.float64)) * DX
    y = (orc.cell_unpack(st["cell"])[1] + st["oy"].astype(np.float64)) * DX
    vsx, vsy = np.zeros(N), np.zeros(N)
    cx, cy = np.zeros(N, int), np.zeros(N, int)
    dist = np.zeros((4, nb), int)
    half = (nb - 1) // 2
    mid = N // 2
    for k in range(len(x)):
        jp, kp = int(x[k] / DX), int(y[k] / DX)   # C truncation toward zero
        vx, vy = float(st["vx"][k]), float(st["vy"][k])
        if jp < N and x[k] > 0.0 and kp == mid:
            vsx[jp] += vx
            cx[jp] += 1
        if jp == mid and kp < N and y[k] > 0.0:
            vsy[kp] += vy
            cy[kp] += 1
        vb = int((vx / vt) * divide)
        for d, (lo, hi) in ((0, (loc[0] - spread, loc[0] + spread)), (2, (loc[2] - spread, loc[2] + spread))):
            if lo <= jp <= hi and x[k] > 0.0 and jp < N and kp == mid and abs(vb) <= half:
                dist[d][vb + half] += 1
        vb = int((vy / vt) * divide)
        for d, (lo, hi) in ((1, (loc[1] - spread, loc[1] + spread)), (3, (loc[3] - spread, loc[3] + spread))):
            if lo <= kp <= hi and y[k] > 0.0 and kp < N and jp == mid and abs(vb) <= half:
                dist[d][vb + half] += 1
    return vsx, vsy, cx, cy, dist


def test_diagnostics_match_the_reference_kernels():
    rng = np.random.default_rng(9)
    n = 40000
    # positions concentrated around the centre row/column, some outside
    jx = np.where(rng.random(n) < 0.5, N // 2, rng.integers(-1, N + 1, n))
    jy = np.where(rng.random(n) < 0.5, N // 2, rng.integers(-1, N + 1, n))
    ox = rng.random(n).astype(np.float32)
    ox[:50] = 0.0                              # x exactly on a node: x > 0 matters at jx = 0
    st = dict(cell=orc.cell_pack(jx, jy), ox=ox, oy=rng.random(n).astype(np.float32),
              vx=(rng.standard_normal(n) * 1e5).astype(np.float32),
              vy=(rng.standard_normal(n) * 1e5).astype(np.float32))
    vt, nb, divide, spread = 1.2e5, 301, 30, 2
    loc = [int(0.9 * N), int(0.5 * N), int(0.1 * N), int(0.5 * N)]
    d = orc.Diag(N, nb)
    orc.lib().orc32_diag_particles(n, orc.P(st["cell"]), orc.P(st["ox"]), orc.P(st["oy"]),
                                   orc.P(st["vx"]), orc.P(st["vy"]), N, vt, nb, divide, spread,
                                   orc.P(np.asarray(loc, np.int32)), orc.P(d.vsx), orc.P(d.vsy),
                                   orc.P(d.cx), orc.P(d.cy), orc.P(d.dist))
    vsx, vsy, cx, cy, dist = reference(st, vt, nb, divide, spread, loc)
    assert np.array_equal(d.cx, cx) and np.array_equal(d.cy, cy)
    assert np.array_equal(d.dist.reshape(4, nb), dist)
    assert dist.sum() > 100
    tol = (cx + 1) * 2.0 ** -20
    assert np.all(np.abs(d.vsx / 1048576.0 - vsx) <= tol)
    assert np.all(np.abs(d.vsy / 1048576.0 - vsy) <= (cy + 1) * 2.0 ** -20)
    # per-step finalisation: mean velocity of the cells that hold particles
    orc.lib().orc_diag_vfinal(N, orc.P(d.vsx), orc.P(d.vsy), orc.P(d.cx), orc.P(d.cy),
                              orc.P(d.vpx), orc.P(d.vpy))
    with np.errstate(invalid="ignore", divide="ignore"):
        want = np.where(cx > 0, vsx / np.maximum(cx, 1), 0.0)
    assert np.allclose(d.vpx, want, rtol=1e-9, atol=1e-6)
    assert not d.cx.any() and not d.vsx.any()
