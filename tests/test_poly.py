"""GMRES polynomial preconditioner coefficients (emerge_amd.fdsolve.poly_coefficients).

Pinned by what the mathematics fixes: after a complete Arnoldi run the
harmonic Ritz values are the eigenvalues; after d < n steps the residual
polynomial prod (1 - z/theta_i) is the GMRES residual polynomial, so applied to
the start vector it gives the GMRES residual norm after d steps; and the
recurrence the GPU kernels evaluate (y = c_0 x, q <- q - c_i B q, y += c_{i+1} q)
gives 1 - z p(z) = prod (1 - c_i z)."""
import numpy as np

from emerge_amd.fdsolve import poly_coefficients


def _arnoldi(A, v, d):
    n = A.shape[0]
    V = np.zeros((n, d + 1), complex)
    H = np.zeros((d + 1, d), complex)
    V[:, 0] = v / np.linalg.norm(v)
    for j in range(d):
        w = A @ V[:, j]
        for _ in range(2):
            for i in range(j + 1):
                h = np.vdot(V[:, i], w)
                H[i, j] += h
                w = w - h * V[:, i]
        H[j + 1, j] = np.linalg.norm(w)
        if H[j + 1, j] > 0:
            V[:, j + 1] = w / H[j + 1, j]
    return H


def _matrix(n, seed):
    rng = np.random.default_rng(seed)
    lam = rng.uniform(0.5, 3, n) + 1j * rng.uniform(-1.5, 1.5, n)
    X = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)) + 4 * np.eye(n)
    return X @ np.diag(lam) @ np.linalg.inv(X), lam, rng


def test_complete_arnoldi_gives_eigenvalues():
    A, lam, rng = _matrix(7, 0)
    H = _arnoldi(A, rng.standard_normal(7) + 1j * rng.standard_normal(7), 7)
    c = poly_coefficients(H)
    theta = 1 / c
    for t in theta:
        assert np.min(np.abs(lam - t)) < 1e-8 * np.abs(lam).max()
    assert abs(abs(theta[0]) - np.abs(theta).max()) < 1e-12   # Leja order starts at the largest


def test_residual_polynomial_is_gmres_optimal():
    A, lam, rng = _matrix(40, 1)
    v = rng.standard_normal(40) + 1j * rng.standard_normal(40)
    for d in (2, 3, 5, 8):
        H = _arnoldi(A, v, d)
        c = poly_coefficients(H)
        r = v.copy()
        for ci in c:
            r = r - ci * (A @ r)
        # GMRES residual after d steps from x0 = 0: min |beta e1 - H y|
        e1 = np.zeros(d + 1, complex)
        e1[0] = np.linalg.norm(v)
        y = np.linalg.lstsq(H, e1, rcond=None)[0]
        gm = np.linalg.norm(e1 - H @ y)
        assert abs(np.linalg.norm(r) - gm) <= 1e-8 * np.linalg.norm(v), (d, np.linalg.norm(r), gm)


def test_kernel_recurrence_matches_product_form():
    rng = np.random.default_rng(2)
    c = rng.standard_normal(6) + 1j * rng.standard_normal(6)
    for z in rng.standard_normal(5) + 1j * rng.standard_normal(5):
        y, q = c[0] * 1.0, 1.0 + 0j
        for i in range(len(c) - 1):
            q = q - c[i] * z * q
            y = y + c[i + 1] * q
        assert abs((1 - z * y) - np.prod(1 - c * z)) < 1e-12 * max(1.0, abs(np.prod(1 - c * z)))


def test_degenerate_returns_none():
    H = np.zeros((3, 2), complex)
    assert poly_coefficients(H) is None
