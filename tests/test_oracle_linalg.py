"""Pins the oracle's dense kernels to closed-form identities:
pinv (Matrix.c:467-541, LAPACK in the reference) by the Moore-Penrose
conditions and numpy; pivoted MGS (Matrix.c:215-307) by orthonormality,
span and the rank cut; the DFTs (FFTW in the reference, FFT.c:10-150) by
numpy's FFT and by direct circular convolution of the padded grids."""
import numpy as np

import oracle


def test_pinv_moore_penrose():
    rng = np.random.default_rng(0)
    for m, n in ((49, 64), (49, 125), (7, 3)):
        X = rng.normal(size=(m, n)) @ np.diag(np.logspace(0, -6, n)) if n <= m else rng.normal(size=(m, n))
        P = oracle.pinv(X)
        np.testing.assert_allclose(X @ P @ X, X, atol=1e-10 * np.abs(X).max())
        np.testing.assert_allclose(P @ X @ P, P, atol=1e-9 * np.abs(P).max())
        np.testing.assert_allclose((X @ P).T, X @ P, atol=1e-10)
        np.testing.assert_allclose((P @ X).T, P @ X, atol=1e-10)
        np.testing.assert_allclose(P, np.linalg.pinv(X), atol=1e-9 * np.abs(P).max())


def test_pivoted_mgs_basis():
    rng = np.random.default_rng(1)
    rows, cols, r = 30, 200, 12
    A = rng.normal(size=(rows, r)) @ rng.normal(size=(r, cols))
    A += 1e-9 * rng.normal(size=A.shape)
    Q = oracle.colbasis_mgs(A, 1e-5)
    assert Q.shape[1] == r
    np.testing.assert_allclose(Q.T @ Q, np.eye(r), atol=1e-12)
    resid = A - Q @ (Q.T @ A)
    assert np.linalg.norm(resid) <= 1e-5 * np.linalg.norm(A)
    # first pivot is the largest column
    j = np.argmax(np.linalg.norm(A, axis=0))
    np.testing.assert_allclose(Q[:, 0], A[:, j] / np.linalg.norm(A[:, j]), atol=1e-14)
    # full-rank matrix keeps every column
    B = rng.normal(size=(10, 6))
    assert oracle.colbasis_mgs(B, 1e-5).shape[1] == 6


def test_dft_matches_numpy():
    rng = np.random.default_rng(2)
    for N in (7, 9, 11):
        g = rng.normal(size=(N, N, N))
        S = oracle.dft_r2c(g)
        np.testing.assert_allclose(S, np.fft.rfftn(g), atol=1e-11)
        back = oracle.dft_c2r(S, N)
        np.testing.assert_allclose(back / N ** 3, g, atol=1e-13)


def test_padded_fft_convolution_is_direct_sum():
    """crop(IFFT(FFT(t) * FFT(pad(g))))/N^3 == sum_b G(h(shift + a - b)) g_b  (FFT.c:10-150)."""
    rng = np.random.default_rng(3)
    for n in (4, 5):
        N = 2 * n - 1
        off = n // 2
        h = 0.37
        shift = np.array([2, -3, 0]) * (n - 1)
        t = np.zeros((N, N, N))
        idx = np.arange(N)
        l = np.where(idx > N // 2, N - idx, -idx)
        for i in range(N):
            for j in range(N):
                for k in range(N):
                    p = h * (shift - np.array([l[i], l[j], l[k]]))
                    t[i, j, k] = 1.0 / np.linalg.norm(p)
        g = rng.normal(size=(n, n, n))
        P = np.zeros((N, N, N))
        P[off:off + n, off:off + n, off:off + n] = g
        out = oracle.dft_c2r(oracle.dft_r2c(t) * oracle.dft_r2c(P), N)[off:off + n, off:off + n, off:off + n] / N ** 3
        grid = np.stack(np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij"), -1).reshape(-1, 3)
        direct = np.zeros(n ** 3)
        for ai, a in enumerate(grid):
            d = h * (shift + a - grid)
            direct[ai] = np.sum(g.reshape(-1) / np.linalg.norm(d, axis=1))
        np.testing.assert_allclose(out.reshape(-1), direct, rtol=1e-12, atol=1e-12)
