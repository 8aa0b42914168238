"""Pins of the oracle's SBP-6 first-derivative operator and metric pipeline.

The operator (JU_xJU_yJU_z6.f90:12-809, metrics.f90:16-488) is fixed by
mathematics, not by retyping its table:
  * summation by parts: a diagonal norm H exists with H D + (H D)^T = B,
    B = diag(-1, 0, ..., 0, 1), and its first entry is the published
    13649/43200 = 0.315949074074074 (block.f90:242 uses 1/this as tau0);
  * accuracy: closure rows differentiate polynomials of degree <= 3 exactly,
    interior rows degree <= 6;
  * the right closure is the antisymmetric mirror of the left one.
A mistyped, transposed or sign-flipped coefficient fails one of these.
"""
import numpy as np
import pytest


def dmatrix(O, n):
    D = np.zeros((n, n))
    for j in range(n):
        e = np.zeros(n)
        e[j] = 1.0
        D[:, j] = O.apply_d1(e, 1.0)
    return D


def test_sbp_identity_and_norm(oracle):
    n = 30
    D = dmatrix(oracle, n)
    # unknown boundary weights h1..h6 (interior weights are 1 with unit spacing)
    rows, rhs = [], []
    for i in range(9):
        for j in range(i, 9):
            coef = np.zeros(6)
            const = 0.0
            # h_i D_ij + h_j D_ji = B_ij
            for (a, b) in ((i, j), (j, i)):
                if a < 6:
                    coef[a] += D[a, b]
                else:
                    const += D[a, b]
            B = -1.0 if (i == 0 and j == 0) else 0.0
            rows.append(coef)
            rhs.append(B - const)
    A = np.array(rows)
    h, res, *_ = np.linalg.lstsq(A, np.array(rhs), rcond=None)
    resid = np.abs(A @ h - np.array(rhs)).max()
    assert resid < 1e-12, resid
    assert abs(h[0] - 13649.0 / 43200.0) < 1e-12
    assert abs(1.0 / 0.315949074074074 - 1.0 / h[0]) < 1e-11
    # full operator: Q = H D satisfies Q + Q^T = B everywhere
    H = np.ones(n)
    H[:6] = h
    H[-6:] = h[::-1]
    Q = H[:, None] * D
    S = Q + Q.T
    Bm = np.zeros((n, n))
    Bm[0, 0], Bm[-1, -1] = -1.0, 1.0
    assert np.abs(S - Bm).max() < 1e-12


def test_polynomial_exactness(oracle):
    n = 40
    h = 1.0 / (n - 1)
    x = np.arange(n) * h
    for k in range(0, 8):
        u = x ** k
        du = oracle.apply_d1(u, h)
        exact = k * x ** (k - 1) if k > 0 else np.zeros(n)
        err = np.abs(du - exact)
        scale = max(1.0, k)
        if k <= 3:
            assert err.max() < 1e-10 * scale, (k, err.max())
        if k <= 6:
            assert err[6:n - 6].max() < 1e-10 * scale, (k, err[6:n - 6].max())
        else:
            assert err[6:n - 6].max() > 1e-8   # degree 7 is not exact: stencil is 6th order


def test_mirror_antisymmetry(oracle):
    n = 25
    D = dmatrix(oracle, n)
    R = -D[::-1, ::-1]
    assert np.array_equal(D, R)


def test_metrics_affine_map_exact(oracle):
    """For an affine map x = A xi + b the metric terms are the exact inverse
    Jacobian matrix (SBP rows are exact for linear functions)."""
    n = (14, 15, 16)
    rng = np.random.default_rng(3)
    A = rng.uniform(-1, 1, (3, 3)) + 3 * np.eye(3)
    b = rng.uniform(-5, 5, 3)
    q = np.linspace(0, 1, n[0]); r = np.linspace(0, 1, n[1]); s = np.linspace(0, 1, n[2])
    S, R, Qq = np.meshgrid(s, r, q, indexing="ij")
    xi = np.stack([Qq, R, S])
    X = np.einsum("ij,jzyx->izyx", A, xi) + b[:, None, None, None]
    metx, mety, metz, J = oracle.metrics(X)
    Ainv = np.linalg.inv(A)          # d xi / d x: row = xi component
    assert np.allclose(J, np.linalg.det(A), rtol=1e-12, atol=0)
    for comp, met in enumerate((metx, mety, metz)):   # x, y, z derivative
        for m in range(3):                            # q, r, s
            assert np.allclose(met[m], Ainv[m, comp], rtol=1e-11, atol=1e-12)


def test_metrics_cubic_stretching_exact(oracle):
    """x = f(q) cubic: closure rows are exact to degree 3, so q_x = 1/f'(q)."""
    n = (20, 13, 13)
    q = np.linspace(0, 1, n[0])
    f = 2.0 + 3.0 * q + 0.5 * q ** 2 - 0.3 * q ** 3
    fp = 3.0 + 1.0 * q - 0.9 * q ** 2
    X = np.zeros((3, n[2], n[1], n[0]))
    X[0] = f[None, None, :]
    X[1] = np.linspace(0, 2, n[1])[None, :, None]
    X[2] = np.linspace(0, 4, n[2])[:, None, None]
    metx, mety, metz, J = oracle.metrics(X)
    assert np.allclose(metx[0], (1.0 / fp)[None, None, :], rtol=1e-11)
    assert np.allclose(J, (fp * 2.0 * 4.0)[None, None, :], rtol=1e-11)
