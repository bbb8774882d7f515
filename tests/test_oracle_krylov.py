"""Pins the oracle's restatement of the Krylov methods besides PCG
(oracle/krylov.py: warp.optim.linear bicgstab / cr / gmres as called from
the reference's solver.py:1219-1252) by properties the mathematics fixes:

- CR's k-th iterate minimises r^T M r over x0 + K_k(MA, M r0);
- one GMRES(k) cycle minimises ||b - A x|| over x0 + M K_k(AM, r0);
- BiCGSTAB, CR and GMRES converge to the dense solution of symmetric
  positive definite and non-symmetric systems, with and without the
  fixed-node projector, and a converged start takes no step;
- point-Jacobi blocks invert the diagonal (1 where it is zero).
"""

import numpy as np
import pytest

import krylov


def spd(n, seed, cond=50.0):
    rng = np.random.default_rng(seed)
    q, _ = np.linalg.qr(rng.normal(size=(n, n)))
    return q @ np.diag(np.geomspace(1.0, cond, n)) @ q.T


def block_spd(n, seed):
    rng = np.random.default_rng(seed)
    M = np.zeros((n, n))
    for i in range(0, n, 3):
        B = rng.normal(size=(3, 3))
        M[i:i + 3, i:i + 3] = B @ B.T + 3 * np.eye(3)
    return M


def krylov_basis(op, v, k):
    cols = [v]
    for _ in range(k - 1):
        cols.append(op @ cols[-1])
    q, _ = np.linalg.qr(np.array(cols).T)
    return q


@pytest.mark.parametrize("k", [1, 2, 3, 5])
def test_cr_iterate_minimises_preconditioned_residual(k):
    n = 24
    A, M = spd(n, 1), block_spd(n, 2)
    rng = np.random.default_rng(3)
    b, x0 = rng.normal(size=n), rng.normal(size=n)
    x, iters, _ = krylov.cr(A, b, x0, M, 1e-30, k)
    assert iters == k
    r0 = b - A @ x0
    Q = krylov_basis(M @ A, M @ r0, k)
    L = np.linalg.cholesky(M)  # r^T M r = ||L^T r||^2
    c, *_ = np.linalg.lstsq(L.T @ A @ Q, L.T @ r0, rcond=None)
    np.testing.assert_allclose(x, x0 + Q @ c, rtol=1e-8, atol=1e-10)


@pytest.mark.parametrize("k", [1, 2, 4, 6])
@pytest.mark.parametrize("symmetric", [True, False])
def test_gmres_cycle_minimises_residual(k, symmetric):
    n = 21
    rng = np.random.default_rng(k)
    A = spd(n, 4) if symmetric else spd(n, 4) + 0.3 * rng.normal(size=(n, n))
    M = block_spd(n, 5)
    b, x0 = rng.normal(size=n), rng.normal(size=n)
    x, iters, _ = krylov.gmres(A, b, x0, M, 1e-30, k, restart=k)
    assert iters == k
    r0 = b - A @ x0
    Q = krylov_basis(A @ M, r0, k)
    c, *_ = np.linalg.lstsq(A @ M @ Q, r0, rcond=None)
    np.testing.assert_allclose(x, x0 + M @ Q @ c, rtol=1e-8, atol=1e-10)


@pytest.mark.parametrize("method", ["bicgstab", "cr", "gmres"])
@pytest.mark.parametrize("symmetric", [True, False])
def test_methods_converge_to_the_dense_solution(method, symmetric):
    if method == "cr" and not symmetric:
        pytest.skip("conjugate residual needs a symmetric operator")
    n = 30
    rng = np.random.default_rng(7)
    A = spd(n, 8, cond=20.0)
    if not symmetric:
        A = A + 0.2 * rng.normal(size=(n, n))
    M = krylov.point_jacobi(A, "diag")
    b = rng.normal(size=n)
    fixed = (1, 4)
    Op = krylov.projected_system(A, fixed)
    bp = b.copy()
    for i in fixed:
        bp[3 * i:3 * i + 3] = 0.0
    x, iters, rr = krylov.METHODS[method](Op, bp, np.zeros(n), M, 1e-12, 300)
    exact = np.linalg.solve(Op, bp)
    np.testing.assert_allclose(x, exact, rtol=1e-9, atol=1e-9)
    assert rr <= krylov.atol_sq(bp, 1e-12) and iters > 0
    # a start that already meets the tolerance takes no step
    x2, it2, _ = krylov.METHODS[method](Op, bp, exact, M, 1e-6, 50)
    np.testing.assert_array_equal(x2, exact)
    assert it2 == 0


def test_point_jacobi_kinds():
    A = np.diag([2.0, -4.0, 0.0]) + 0.1
    np.testing.assert_allclose(np.diag(krylov.point_jacobi(A, "diag")), [1 / 2.1, 1 / -3.9, 1 / 0.1])
    A = np.diag([2.0, -4.0, 0.0])
    np.testing.assert_allclose(np.diag(krylov.point_jacobi(A, "diag")), [0.5, -0.25, 1.0])
    np.testing.assert_allclose(np.diag(krylov.point_jacobi(A, "diag_abs")), [0.5, 0.25, 1.0])
    np.testing.assert_array_equal(krylov.point_jacobi(A, "id"), np.eye(3))
