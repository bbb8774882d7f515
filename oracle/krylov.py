"""FP64 restatement of the step's Krylov methods besides PCG (test infrastructure).

The reference's SolverFEM._make_linear_solver (solver.py:1219-1252) calls
warp.optim.linear.bicgstab / cr / gmres (warp-lang 1.17, the reference's
pinned dependency) with tol = linear_solver_tolerance, atol = 0, maxiter, the
preconditioner M and the warm start; the preconditioner types "diag",
"diag_abs" and "id" come from warp.optim.linear.preconditioner
(solver.py:1155-1156).  These functions restate those methods on dense
matrices, with every update masked by the latest residual test
rr > atol^2 (atol = max(tol ||b||, eps32^(9/4))), so that running the whole
budget gives the result of stopping at convergence.  Only ``tests/`` import
this module, as the checker.

Parity unpinned: Warp's exact bicgstab / cr / gmres iterate sequences and
their breakdown guards (warp-lang cannot be imported here); these
restatements are pinned through the minimisation properties of the cr and
gmres iterates, convergence to dense solves and the point-Jacobi kinds
(tests/test_oracle_krylov.py).
"""

from __future__ import annotations

import numpy as np

EPS32 = float(np.finfo(np.float32).eps)


def atol_sq(b, tol):
    atol = max(tol * float(np.linalg.norm(b)), 0.0, EPS32 ** 2.25)
    return atol * atol


def point_jacobi(A, kind="diag"):
    """Inverse of the diagonal entries (of their magnitudes for "diag_abs"),
    1 where zero; the identity for "id"."""
    d = np.diag(A).astype(np.float64).copy()
    if kind == "id":
        return np.eye(len(d))
    if kind == "diag_abs":
        d = np.abs(d)
    inv = np.where(d == 0.0, 1.0, 1.0 / np.where(d == 0.0, 1.0, d))
    return np.diag(inv)


def bicgstab(A, b, x0, M, tol, maxiter):
    """Right-preconditioned BiCGSTAB; returns (x, live iterations, rr)."""
    x = np.array(x0, dtype=np.float64)
    r = b - A @ x
    rr, a2 = r @ r, atol_sq(b, tol)
    r0, p, rho = r.copy(), r.copy(), r @ r
    v = np.zeros_like(r)
    beta = omega = 0.0
    iters = 0
    for it in range(maxiter):
        if it > 0:
            p = r + beta * (p - omega * v)
        y = M @ p
        v = A @ y
        r0v = r0 @ v
        live = rr > a2
        iters += int(live)
        alpha = rho / r0v if live and r0v != 0.0 else 0.0
        x = x + alpha * y
        r = r - alpha * v
        rr = r @ r
        z = M @ r
        t = A @ z
        st, tt = r @ t, t @ t
        omega = st / tt if rr > a2 and tt > 0.0 else 0.0
        x = x + omega * z
        r = r - omega * t
        rr = r @ r
        rho_new = r0 @ r
        beta = (rho_new / r0v) / omega if rr > a2 and r0v != 0.0 and omega != 0.0 else 0.0
        rho = rho_new
    return x, iters, rr


def cr(A, b, x0, M, tol, maxiter):
    """Preconditioned conjugate residual; returns (x, live iterations, rr)."""
    x = np.array(x0, dtype=np.float64)
    r = b - A @ x
    rr, a2 = r @ r, atol_sq(b, tol)
    z = M @ r
    az = A @ z
    zaz = z @ az
    p, ap = z.copy(), az.copy()
    beta = 0.0
    iters = 0
    for it in range(maxiter):
        if it > 0:
            p = z + beta * p
            ap = az + beta * ap
        y = M @ ap
        yap = ap @ y
        live = rr > a2
        iters += int(live)
        alpha = zaz / yap if live and yap > 0.0 else 0.0
        x = x + alpha * p
        r = r - alpha * ap
        z = z - alpha * y
        rr = r @ r
        az = A @ z
        zaz_new = z @ az
        beta = zaz_new / zaz if rr > a2 and zaz > 0.0 else 0.0
        zaz = zaz_new
    return x, iters, rr


def gmres(A, b, x0, M, tol, maxiter, restart=31):
    """Restarted right-preconditioned GMRES (classical Gram-Schmidt Arnoldi,
    Givens least squares with pivot tolerance eps32^2); returns (x, live
    iterations counted per cycle of `restart`, rr)."""
    k = max(1, min(restart, maxiter))
    x = np.array(x0, dtype=np.float64)
    r = b - A @ x
    rr, a2 = r @ r, atol_sq(b, tol)
    iters = 0
    for _ in range((maxiter + k - 1) // k):
        if not rr > a2:
            continue
        iters += k
        beta = np.sqrt(rr)
        V = np.zeros((k + 1, len(b)))
        H = np.zeros((k + 1, k))
        V[0] = r / beta
        for j in range(k):
            w = A @ (M @ V[j])
            H[: j + 1, j] = V[: j + 1] @ w
            w = w - H[: j + 1, j] @ V[: j + 1]
            H[j + 1, j] = np.sqrt(w @ w)
            V[j + 1] = w / H[j + 1, j] if H[j + 1, j] != 0.0 else w
        y = np.zeros(k)
        g = beta
        kmax = k
        R = H.copy()
        for i in range(k):
            ha, hb = R[i, i], R[i + 1, i]
            nn = ha * ha + hb * hb
            if nn < EPS32 ** 2:
                kmax = i
                break
            n = np.sqrt(nn)
            c, s = ha / n, hb / n
            top, bot = R[i, i:].copy(), R[i + 1, i:].copy()
            R[i, i:] = c * top + s * bot
            R[i + 1, i:] = c * bot - s * top
            y[i] = c * g
            g = -s * g
        for i in range(kmax - 1, -1, -1):
            y[i] = (y[i] - R[i, i + 1:kmax] @ y[i + 1:kmax]) / R[i, i]
        y[kmax:] = 0.0
        x = x + M @ (V[:k].T @ y)
        r = b - A @ x
        rr = r @ r
    return x, iters, rr


METHODS = {"bicgstab": bicgstab, "cr": cr, "gmres": gmres}


def projected_system(A, fixed):
    """Operator P A P + I - P of the fixed-node projector (fixed_particles.py:29-53)
    for a dense 3P x 3P matrix and node indices `fixed`."""
    n = A.shape[0]
    free = np.ones(n, bool)
    for i in fixed:
        free[3 * i:3 * i + 3] = False
    Op = A * np.outer(free, free)
    idx = np.nonzero(~free)[0]
    Op[idx, idx] = 1.0
    return Op
