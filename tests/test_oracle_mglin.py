"""Oracle multigrid (orc_mglin, orc_mg_tol) pinned three ways:
bit-for-bit against the reference's own Mglin (mglin.cpp compiled into
oracle/_ref), against the exact discrete eigen-solution of the 5-point
Laplacian, and against a dense direct solve on small grids."""
import numpy as np
import pytest

import orc


def laplacian_dense(n, length):
    """5-point Laplacian on the (n-2)^2 interior nodes, zero Dirichlet walls."""
    h = length / (n - 1)
    m = n - 2
    A = np.zeros((m * m, m * m))
    for i in range(m):
        for j in range(m):
            k = i * m + j
            A[k, k] = -4.0
            for di, dj in ((1, 0), (-1, 0), (0, 1), (0, -1)):
                a, b = i + di, j + dj
                if 0 <= a < m and 0 <= b < m:
                    A[k, a * m + b] = 1.0
    return A / (h * h)


@pytest.mark.skipif(not orc.ref_mglin_available(), reason="oracle/_ref not built")
@pytest.mark.parametrize("n", [3, 5, 9, 17, 33, 65, 129, 257])
@pytest.mark.parametrize("ncycle", [1, 2, 3])
def test_bit_exact_against_reference_mglin(n, ncycle):
    rng = np.random.default_rng(100 * n + ncycle)
    f = rng.standard_normal((n, n)) * 1e4
    L = 0.01 * (1 + ncycle)
    a = orc.mglin(f, ncycle, L)
    b = orc.ref_mglin(f, ncycle, L)
    assert np.array_equal(a, b)


@pytest.mark.parametrize("n,p,q", [(33, 1, 1), (65, 2, 3), (129, 5, 1)])
def test_discrete_eigenmode(n, p, q):
    L = 0.02
    h = L / (n - 1)
    x = np.arange(n) * h
    f = np.outer(np.sin(p * np.pi * x / L), np.sin(q * np.pi * x / L))
    lam = -(4.0 / h ** 2) * (np.sin(p * np.pi * h / (2 * L)) ** 2 + np.sin(q * np.pi * h / (2 * L)) ** 2)
    exact = f / lam
    u, cycles, rr = orc.mg_tol(f, L, 1e-13)
    assert rr <= 1e-13
    assert np.max(np.abs(u - exact)) <= 1e-9 * np.max(np.abs(exact))
    # fixed-work FMG (ncycle=2) must already be close to the discrete solution
    u2 = orc.mglin(f, 2, L)
    assert np.max(np.abs(u2 - exact)) <= 2e-2 * np.max(np.abs(exact))


@pytest.mark.parametrize("n", [9, 17])
def test_dense_direct_solve(n):
    L = 0.37
    rng = np.random.default_rng(n)
    f = rng.standard_normal((n, n))
    A = laplacian_dense(n, L)
    u_dense = np.linalg.solve(A, f[1:-1, 1:-1].ravel()).reshape(n - 2, n - 2)
    u, _, _ = orc.mg_tol(f, L, 1e-14)
    assert np.allclose(u[1:-1, 1:-1], u_dense, rtol=0, atol=1e-10 * np.abs(u_dense).max())
    assert np.all(u[0, :] == 0) and np.all(u[-1, :] == 0) and np.all(u[:, 0] == 0)
    # FMG + more fixed cycles converges to the same solution
    u30 = orc.mglin(f, 30, L)
    assert np.allclose(u30[1:-1, 1:-1], u_dense, rtol=0, atol=1e-9 * np.abs(u_dense).max())


def test_grid_size_rejected():
    with pytest.raises(ValueError):
        orc.mglin(np.zeros((6, 6)), 2, 1.0)
