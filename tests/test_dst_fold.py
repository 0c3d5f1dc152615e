"""The algebra of the folded 64 x 64 sine transform that the coarse DST solve
uses on the device (k_fields.hip, dstf_solve): the row and column folds,
the odd- and even-frequency half tables, the eigenvalue scale in split order
and the unfolds reproduce the full transform S (D o (S X S)) S, and that
transform solves the 5-point Poisson problem on the 63 x 63 interior.  A
dropped sign in a fold, a swapped half or a misplaced split index fails the
first comparison; a wrong table or scale fails the second."""
import numpy as np

M = 64


def tables():
    a = np.arange(1, M)
    S = np.zeros((M, M))
    S[:M - 1, :M - 1] = np.sin(np.pi * (np.outer(a, a) % (2 * M)) / M)
    lam = -4.0 * np.sin(np.pi * a / (2.0 * M)) ** 2
    D = np.zeros((M, M))
    D[:M - 1, :M - 1] = (2.0 / M) ** 2 / (lam[:, None] + lam[None, :])
    i, p = np.arange(32)[:, None], np.arange(32)[None, :]
    H1 = np.sin(np.pi * (((2 * i + 1) * (p + 1)) % (2 * M)) / M)
    H2 = np.sin(np.pi * (((2 * i + 2) * (p + 1)) % (2 * M)) / M)
    H2[31, :] = 0.0
    H2[:, 31] = 0.0
    # split index -> natural frequency index: odd frequencies first
    nat = np.array([2 * r for r in range(32)] + [2 * r + 1 for r in range(31)] + [-1])
    ok = nat >= 0
    Ds = np.zeros((M, M))
    Ds[np.ix_(ok, ok)] = D[np.ix_(nat[ok], nat[ok])]
    return S, D, H1, H2, Ds


def fold_rows(X):
    W = np.zeros_like(X)
    o = np.arange(31)
    W[o] = X[o] + X[62 - o]
    W[31] = X[31]
    W[32 + o] = X[o] - X[62 - o]
    return W


def unfold_rows(W):
    X = np.zeros_like(W)
    o = np.arange(31)
    X[o] = W[o] + W[32 + o]
    X[31] = W[31]
    X[62 - o] = W[o] - W[32 + o]
    return X


def halves(B, H1, H2, transpose):
    C = np.zeros_like(B)
    for half, H in ((0, H1), (1, H2)):
        C[32 * half:32 * half + 32] = (H.T if transpose else H) @ B[32 * half:32 * half + 32]
    return C


def folded_solve(X, H1, H2, Ds, h2):
    Y = halves(fold_rows(X), H1, H2, False)          # rows to split order
    Z = halves(fold_rows(Y.T), H1, H2, False).T      # columns too
    Z = Z * (Ds * h2)
    U = unfold_rows(halves(Z, H1, H2, True))         # rows back to natural order
    return unfold_rows(halves(U.T, H1, H2, True)).T  # columns too


def test_folded_transform_matches_full_transform_and_solves_poisson():
    S, D, H1, H2, Ds = tables()
    rng = np.random.default_rng(7)
    X = np.zeros((M, M))
    X[:63, :63] = rng.standard_normal((63, 63))
    h2 = 0.37
    full = S @ ((S @ X @ S) * (D * h2)) @ S
    got = folded_solve(X, H1, H2, Ds, h2)
    assert np.abs(got - full).max() <= 1e-13 * np.abs(full).max()
    assert not got[63].any() and not got[:, 63].any()
    # the 5-point operator of the solution returns the right-hand side
    u = np.pad(full[:63, :63], 1)
    lap = (u[2:, 1:-1] + u[:-2, 1:-1] + u[1:-1, 2:] + u[1:-1, :-2] - 4 * u[1:-1, 1:-1]) / h2
    assert np.abs(lap - X[:63, :63]).max() <= 1e-12 * np.abs(X).max()
