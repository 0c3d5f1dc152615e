"""Oracle field functions: E stencil (fields.cpp:230-268), potential with
Dirichlet folding (fields.cpp:169-220) and the fixed-point charge -> rhs path,
pinned by closed forms and a dense solve."""
import math

import numpy as np

import orc
from test_oracle_mglin import laplacian_dense


def test_linear_phi_gives_constant_field():
    N, L = 33, 0.01
    dx = L / (N - 1)
    x = np.arange(N) * dx
    a, b = 120.0, -45.0
    phi = a * x[None, :] + b * x[:, None]        # phi[j][i] = a x_i + b y_j
    ex, ey = orc.ref_efield(phi, dx)
    ex = ex[: N * N].reshape(N, N)
    ey = ey[: N * N].reshape(N, N)
    # interior rows: Ex = -a everywhere incl. the extrapolated wall columns
    assert np.allclose(ex[1:-1, :], -a, rtol=1e-12)
    assert np.all(ex[0, :] == 0) and np.all(ex[-1, :] == 0)      # never written
    assert np.allclose(ey[:, 1:-1], -b, rtol=1e-12)
    assert np.all(ey[:, 0] == 0) and np.all(ey[:, -1] == 0)
    # packed per-node restatement == sequential statement, bit for bit
    ex2, ey2, e32 = orc.efield_pack(phi, dx)
    assert np.array_equal(ex2, ex) and np.array_equal(ey2, ey)
    assert np.array_equal(e32[0:2 * N * N:2], ex.ravel().astype(np.float32))
    assert np.all(e32[2 * N * N:] == 0)


def test_efield_pack_matches_reference_statement_random():
    N, L = 65, 0.01
    dx = L / (N - 1)
    phi = np.random.default_rng(1).standard_normal((N, N))
    ex, ey = orc.ref_efield(phi, dx)
    ex2, ey2, _ = orc.efield_pack(phi, dx)
    assert np.array_equal(ex2.ravel(), ex[: N * N]) and np.array_equal(ey2.ravel(), ey[: N * N])
    assert np.all(ex[N * N:] == 0) and np.all(ey[N * N:] == 0)


def test_potential_with_dirichlet_walls_against_dense_solve():
    N, L, eps = 17, 0.05, 8.854e-12
    dx = L / (N - 1)
    rng = np.random.default_rng(5)
    rho = rng.standard_normal((N, N)) * 1e-9
    pl, pr, pb, pt = 3.0, -2.0, 1.5, 0.5
    phi = np.zeros(N * N)
    rc = orc.lib().orc_ref_grid_potential(orc.P(np.ascontiguousarray(rho)), N, dx, dx, eps, pl, pr,
                                          pb, pt, 0.0, 0.0, 40, L, orc.P(phi))
    assert rc == 0
    phi = phi.reshape(N, N)
    # dense: interior Laplacian with the wall values moved to the right-hand side
    bc = np.zeros((N, N))
    bc[:, 0], bc[:, -1] = pl, pr
    bc[0, :], bc[-1, :] = pb, pt                  # rows win the corners (fields.cpp:209-219)
    A = laplacian_dense(N, L)
    f = (-rho / eps)[1:-1, 1:-1].copy()
    f[:, 0] -= bc[1:-1, 0] / dx ** 2
    f[:, -1] -= bc[1:-1, -1] / dx ** 2
    f[0, :] -= bc[0, 1:-1] / dx ** 2
    f[-1, :] -= bc[-1, 1:-1] / dx ** 2
    u = np.linalg.solve(A, f.ravel()).reshape(N - 2, N - 2)
    assert np.allclose(phi[1:-1, 1:-1], u, rtol=0, atol=1e-8 * np.abs(u).max())
    assert np.all(phi[0, :] == pb) and np.all(phi[-1, :] == pt)
    assert np.all(phi[1:-1, 0] == pl) and np.all(phi[1:-1, -1] == pr)


def test_fixed_to_rhs_postscale_and_sum_order():
    N = 9
    qi = np.full(N * N, 1 << orc.FX_TOTAL_BITS, dtype=np.int64)
    qe = np.full(N * N, 3 << orc.FX_TOTAL_BITS, dtype=np.int64)
    rho = np.zeros(N * N)
    rhs = np.zeros(N * N)
    eps, s_i, s_e = 2.0, 1.0, -1.0
    fx = orc.FX_TOTAL_BITS
    orc.lib().orc_fixed_to_rhs(orc.P(qi), orc.P(qe), N, math.ldexp(s_i, -fx), math.ldexp(s_e, -fx),
                               1, eps, 0.0, 0.0, 0.0, 0.0, 1.0, orc.P(rho), orc.P(rhs))
    rho = rho.reshape(N, N)
    # interior 1 - 3 = -2; edges x2; corners x8 (particle.cpp:830-845)
    assert np.all(rho[1:-1, 1:-1] == -2.0)
    assert np.all(rho[0, 1:-1] == -4.0) and np.all(rho[1:-1, -1] == -4.0)
    assert rho[0, 0] == -16.0 and rho[-1, -1] == -16.0
    # same thing through the reference's FP64 sequence
    q_i = np.full((N, N), 1.0)
    q_e = np.full((N, N), -3.0)
    orc.lib().orc_ref_charge_postscale(orc.P(q_i), N, N)
    orc.lib().orc_ref_charge_postscale(orc.P(q_e), N, N)
    r2 = np.zeros((N, N))
    orc.lib().orc_ref_grid_rho(orc.P(q_i), orc.P(q_e), N, orc.P(r2))
    assert np.array_equal(rho, r2)
    assert np.all(rhs.reshape(N, N) == -(1.0 / eps) * rho)
