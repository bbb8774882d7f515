"""ctypes wrapper of the CPU restatement in ``fem_oracle.cpp``.

Test infrastructure only: imported by ``tests/``, ``__graft_entry__.smoke``
and the CPU-baseline leg of ``bench.py`` as the checker / baseline, never by
the shipped package.  ``bits=64`` is the FP64 oracle, ``bits=32`` mirrors the
reference's fp32 CPU path (nvidia-beer/newton-solver-fem ``solver.py``
``SolverFEM.step`` / ``refresh_tangent``).
"""

from __future__ import annotations

import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liborc_fem.so"

TERM_TET, TERM_HEX, TERM_GROUND, TERM_GRAVITY, TERM_DAMPING, TERM_SPRING, TERM_TRI = 1, 2, 4, 8, 16, 32, 64

_lib = None


def build(force: bool = False) -> Path:
    if force or not LIB_PATH.exists() or LIB_PATH.stat().st_mtime < (HERE / "fem_oracle.cpp").stat().st_mtime:
        subprocess.run(["make", "-s", "-C", str(HERE), "-B" if force else "all"], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        _lib = ctypes.CDLL(str(LIB_PATH))
        d, i = ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)
        _lib.orc_create.restype = ctypes.c_void_p
        _lib.orc_create.argtypes = [ctypes.c_int, ctypes.c_int, d, d, d, d, i, ctypes.c_int, i, d, d, ctypes.c_int, i,
                                    d, d, d, ctypes.c_int, i]
        _lib.orc_destroy.argtypes = [ctypes.c_void_p]
        _lib.orc_set_params.argtypes = [ctypes.c_void_p, d]
        _lib.orc_set_threads.argtypes = [ctypes.c_void_p, ctypes.c_int]
        _lib.orc_construct.argtypes = [ctypes.c_void_p, ctypes.c_double]
        _lib.orc_refresh.argtypes = [ctypes.c_void_p, d, ctypes.c_double]
        _lib.orc_nnz.argtypes = [ctypes.c_void_p]
        _lib.orc_nnz.restype = ctypes.c_int
        _lib.orc_get_matrix.argtypes = [ctypes.c_void_p, i, i, d, d]
        _lib.orc_get_dv.argtypes = [ctypes.c_void_p, d]
        _lib.orc_set_dv.argtypes = [ctypes.c_void_p, d]
        _lib.orc_step.argtypes = [ctypes.c_void_p, d, d, d, ctypes.c_void_p, ctypes.c_void_p, d, d, ctypes.c_double]
        _lib.orc_step.restype = ctypes.c_int
        _lib.orc_last_stats.argtypes = [ctypes.c_void_p, d]
        _lib.orc_eval_terms.argtypes = [ctypes.c_void_p, ctypes.c_int, d, d, ctypes.c_void_p, d]
        _lib.orc_clamp_stretch.argtypes = [ctypes.c_int, d, ctypes.c_double, ctypes.c_double, d]
        _lib.orc_svd3.argtypes = [d, d, d, d]
        _lib.orc_block_inverse.argtypes = [ctypes.c_int, d, d]
        vp = ctypes.c_void_p
        _lib.orc_vjp_terms.argtypes = [vp, ctypes.c_int, d, d, vp, d, d, d, d, d]
        _lib.orc_diff_forward.argtypes = [vp, d, d, vp, vp, ctypes.c_double, ctypes.c_double, d, d, d]
        _lib.orc_diff_backward.argtypes = [vp, d, d, vp, d, d, d, ctypes.c_double, ctypes.c_double, d, d, d, d, d]
        _lib.orc_set_tet_materials.argtypes = [vp, d]
        _lib.orc_set_springs_tris.argtypes = [vp, ctypes.c_int, i, d, d, d, ctypes.c_int, i, d, d]
    return _lib


def _d(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _i(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def clamp_stretch(F, min_s, max_s, bits=64):
    F, pf = _d(np.asarray(F).reshape(9))
    out, po = _d(np.zeros(9))
    lib().orc_clamp_stretch(bits, pf, float(min_s), float(max_s), po)
    return out.reshape(3, 3)


def svd3(F):
    F, pf = _d(np.asarray(F).reshape(9))
    U, pu = _d(np.zeros(9))
    S, ps = _d(np.zeros(3))
    V, pv = _d(np.zeros(9))
    lib().orc_svd3(pf, pu, ps, pv)
    return U.reshape(3, 3), S, V.reshape(3, 3)


def block_inverse(B, bits=64):
    B, pb = _d(np.asarray(B).reshape(9))
    out, po = _d(np.zeros(9))
    lib().orc_block_inverse(bits, pb, po)
    return out.reshape(3, 3)


def particle_contact_forces(q, qd, radius, flags, ke, kd, kf, mu, cohesion=0.0):
    """Particle-particle contact force on every particle, FP64, all pairs.

    Restates eval_particle_forces (kernels.py:1069-1114) with particle_force
    (kernels.py:1043-1066): for active i and every other active j at distance
    d >= 1e-10 with gap err = d - r_i - r_j <= cohesion, n = (x_i - x_j)/d,
    relative velocity v = v_i - v_j: fn = err ke + min(n.v, 0) kd, tangential
    vt = v - n (n.v), ft = min(|vt| kf, mu |fn|), f_i += -n fn - vt/|vt| ft.
    The hash grid only selects candidate pairs (every pair that can satisfy
    the gap test lies in neighbouring cells), so an all-pairs sum in ascending
    j is the same set of terms.
    """
    q = np.asarray(q, dtype=np.float64).reshape(-1, 3)
    qd = np.asarray(qd, dtype=np.float64).reshape(-1, 3)
    r = np.asarray(radius, dtype=np.float64)
    on = (np.asarray(flags) & 1) != 0
    out = np.zeros_like(q)
    for i in np.nonzero(on)[0]:
        n = q[i] - q
        d = np.sqrt(np.sum(n * n, axis=1))
        err = d - r[i] - r
        sel = on & (np.arange(len(q)) != i) & (d >= 1e-10) & (err <= cohesion)
        if not sel.any():
            continue
        n = n[sel] / d[sel, None]
        v = qd[i] - qd[sel]
        vn = np.sum(n * v, axis=1)
        fn = err[sel] * ke + np.minimum(vn, 0.0) * kd
        vt = v - n * vn[:, None]
        vs = np.sqrt(np.sum(vt * vt, axis=1))
        vt = np.where(vs[:, None] > 0, vt / np.where(vs > 0, vs, 1.0)[:, None], vt)
        ft = np.minimum(vs * kf, mu * np.abs(fn))
        out[i] = np.sum(-n * fn[:, None] - vt * ft[:, None], axis=0)
    return out


def soft_contact_forces(q, qd, radius, count, particle, body_pos, body_vel, normal, ke, kd, kf, mu):
    """Particle-shape soft-contact force on every particle, FP64.

    Restates eval_soft_contacts (kernels.py:1199-1252): for every contact
    k < count with particle p = particle[k] >= 0, penetration
    pen = r_p - (x_p - body_pos[k]).n (n = normal[k], body -> particle); when
    pen > 0: relative velocity rel = v_p - body_vel[k], vn = rel.n,
    vt = rel - vn n, fn = ke pen - kd min(vn, 0); if |vt| > 1e-6 the force is
    fn n - min(kf |vt|, mu max(0, fn)) vt/|vt|, else fn n; forces of one
    particle are summed in contact order.
    """
    q = np.asarray(q, dtype=np.float64).reshape(-1, 3)
    qd = np.asarray(qd, dtype=np.float64).reshape(-1, 3)
    r = np.asarray(radius, dtype=np.float64)
    particle = np.asarray(particle)
    bp = np.asarray(body_pos, dtype=np.float64).reshape(-1, 3)
    bv = np.asarray(body_vel, dtype=np.float64).reshape(-1, 3)
    nrm = np.asarray(normal, dtype=np.float64).reshape(-1, 3)
    out = np.zeros_like(q)
    for k in range(min(int(count), len(particle))):
        p = int(particle[k])
        if p < 0 or p >= len(q):
            continue
        n = nrm[k]
        pen = r[p] - np.dot(q[p] - bp[k], n)
        if pen <= 0.0:
            continue
        rel = qd[p] - bv[k]
        vn = np.dot(rel, n)
        vt = rel - vn * n
        fn = ke * pen - kd * min(vn, 0.0)
        vt_mag = np.sqrt(np.dot(vt, vt))
        if vt_mag > 1.0e-6:
            ft = min(kf * vt_mag, mu * max(0.0, fn))
            out[p] += fn * n - ft * (vt / vt_mag)
        else:
            out[p] += fn * n
    return out


def spring_torque_forces(q, qd, spring_indices, rest_direction, kt, kw):
    """Bend-torque part of eval_springs_linear_and_torque (kernels.py:222-292), FP64.

    Per spring (i, j) with d = (x_i - x_j)/L and d0 = -rest_direction (no
    torque when |d0| < 0.5, L < 1e-6 or |d0 x d| < 1e-6): theta = acos(d.d0),
    axis = (d0 x d)/|d0 x d|, omega = (d x d_dot).axis with
    d_dot = (v_ij - d (d.v_ij))/L, tau = -(kt theta + kw omega) L and
    F = (axis tau) x d / L, applied as -F on i and +F on j.
    """
    q = np.asarray(q, dtype=np.float64).reshape(-1, 3)
    qd = np.asarray(qd, dtype=np.float64).reshape(-1, 3)
    idx = np.asarray(spring_indices).reshape(-1, 2)
    rd = np.asarray(rest_direction, dtype=np.float64).reshape(-1, 3)
    out = np.zeros_like(q)
    for s, (i, j) in enumerate(idx):
        if i < 0 or j < 0:
            continue
        xij, vij = q[i] - q[j], qd[i] - qd[j]
        L = np.linalg.norm(xij)
        if L < 1e-6:
            continue
        d = xij / L
        d0 = -rd[s]
        if np.linalg.norm(d0) < 0.5:
            continue
        theta = np.arccos(np.clip(d @ d0, -1.0, 1.0))
        axis = np.cross(d0, d)
        rl = np.linalg.norm(axis)
        if rl < 1e-6:
            continue
        axis /= rl
        d_dot = (vij - d * (d @ vij)) / L
        omega = np.cross(d, d_dot) @ axis
        tau = -(kt * theta + kw * omega) * L
        F = np.cross(axis * tau, d) / L
        out[i] -= F
        out[j] += F
    return out


class OracleSolver:
    """CPU restatement of ``SolverFEM`` for one model.

    ``host`` is the dict of host arrays a model carries (``model.host``).
    Parameters mirror the reference constructor (solver.py:212-240).
    """

    def __init__(self, host: dict, *, bits=64, dt=1.0 / 60.0, gravity=(0.0, 0.0, -9.81), fixed=(),
                 min_stretch=0.05, max_stretch=20.0, ground_plane=None, ground_ke=1.0e5, ground_kd=1.0e2,
                 ground_kf=1.0e3, ground_mu=0.5, linear_damping=0.0, enable_tet_activations=False,
                 tolerance=1.0e-2, maxiter=50):
        L = lib()
        self.bits = bits
        self.P = len(host["particle_q"])
        self.T = len(host["tet_indices"])
        self.H = len(host["hex_indices"])
        self._keep = []

        def D(a):
            arr, p = _d(a)
            self._keep.append(arr)
            return p

        def I(a):
            arr, p = _i(a)
            self._keep.append(arr)
            return p

        fixed = np.unique(np.asarray(fixed, dtype=np.int64)).astype(np.int32)
        self.fixed = fixed
        self.h = L.orc_create(
            bits, self.P, D(host["particle_q"]), D(host["particle_mass"]), D(host["particle_inv_mass"]),
            D(host["particle_radius"]), I(host["particle_flags"]), self.T, I(host["tet_indices"]),
            D(host["tet_poses"]), D(host["tet_materials"]), self.H, I(host["hex_indices"]), D(host["hex_inv_J0"]),
            D(host["hex_det_J0_w"]), D(host["hex_materials"]), len(fixed), I(fixed),
        )
        self.S = len(host.get("spring_indices", ()))
        self.R = len(host.get("tri_indices", ()))
        if self.S or self.R:
            L.orc_set_springs_tris(
                self.h, self.S, I(host["spring_indices"]), D(host["spring_rest_length"]), D(host["spring_stiffness"]),
                D(host["spring_damping"]), self.R, I(host["tri_indices"]), D(host["tri_poses"]),
                D(host["tri_materials"]),
            )
        self.params = dict(min_stretch=min_stretch, max_stretch=max_stretch, ground_plane=ground_plane,
                           ground_ke=ground_ke, ground_kd=ground_kd, ground_kf=ground_kf, ground_mu=ground_mu,
                           linear_damping=linear_damping, gravity=gravity,
                           enable_tet_activations=enable_tet_activations, tolerance=tolerance, maxiter=maxiter)
        self._push_params()
        self.dt = float(dt)
        L.orc_construct(self.h, float(dt))

    def _push_params(self):
        p = self.params
        g = p["ground_plane"]
        arr = np.array(
            [p["min_stretch"], p["max_stretch"], 0.0 if g is None else 1.0, *(g if g is not None else (0, 0, 1, 0)),
             p["ground_ke"], p["ground_kd"], p["ground_kf"], p["ground_mu"], p["linear_damping"], *p["gravity"],
             1.0 if p["enable_tet_activations"] else 0.0, p["tolerance"], p["maxiter"]],
            dtype=np.float64,
        )
        _, ptr = _d(arr)
        self._params_arr = arr
        lib().orc_set_params(self.h, arr.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))

    def set_threads(self, n: int):
        """Host threads of the CG loops (1: serial order; more only for timing)."""
        lib().orc_set_threads(self.h, int(n))

    def set(self, **kw):
        self.params.update(kw)
        self._push_params()

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.orc_destroy(self.h)
            self.h = None

    def refresh_tangent(self, q: np.ndarray, dt=None) -> np.ndarray:
        dt = self.dt if dt is None else float(dt)
        q = np.ascontiguousarray(q, dtype=np.float64).copy()
        lib().orc_refresh(self.h, q.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), dt)
        return q

    def matrix(self):
        n = lib().orc_nnz(self.h)
        off = np.zeros(self.P + 1, np.int32)
        cols = np.zeros(n, np.int32)
        vals = np.zeros((n, 3, 3))
        minv = np.zeros((self.P, 3, 3))
        lib().orc_get_matrix(self.h, off.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                             cols.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                             vals.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                             minv.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        return off, cols, vals, minv

    def dense(self) -> np.ndarray:
        off, cols, vals, _ = self.matrix()
        out = np.zeros((3 * self.P, 3 * self.P))
        for r in range(self.P):
            for k in range(off[r], off[r + 1]):
                c = cols[k]
                out[3 * r : 3 * r + 3, 3 * c : 3 * c + 3] += vals[k]
        return out

    @property
    def dv(self) -> np.ndarray:
        out = np.zeros((self.P, 3))
        lib().orc_get_dv(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        return out

    @dv.setter
    def dv(self, value):
        v = np.ascontiguousarray(np.broadcast_to(value, (self.P, 3)), dtype=np.float64)
        lib().orc_set_dv(self.h, v.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))

    def step(self, q, qd, dt=None, external_forces=None, activations=None):
        """One substep; returns dict(q_in, qd_in (after fixed restore), f, q, qd, iters)."""
        dt = self.dt if dt is None else float(dt)
        P = self.P
        q = np.ascontiguousarray(q, dtype=np.float64).reshape(P, 3).copy()
        qd = np.ascontiguousarray(qd, dtype=np.float64).reshape(P, 3).copy()
        f = np.zeros((P, 3))
        qo = np.zeros((P, 3))
        qdo = np.zeros((P, 3))
        ext = None if external_forces is None else np.ascontiguousarray(external_forces, dtype=np.float64)
        act = None if activations is None else np.ascontiguousarray(activations, dtype=np.float64)
        dp = ctypes.POINTER(ctypes.c_double)
        it = lib().orc_step(self.h, q.ctypes.data_as(dp), qd.ctypes.data_as(dp), f.ctypes.data_as(dp),
                            None if ext is None else ext.ctypes.data_as(ctypes.c_void_p),
                            None if act is None else act.ctypes.data_as(ctypes.c_void_p),
                            qo.ctypes.data_as(dp), qdo.ctypes.data_as(dp), dt)
        return {"q_in": q, "qd_in": qd, "f": f, "q": qo, "qd": qdo, "iters": it}

    def last_stats(self):
        out = np.zeros(4)
        lib().orc_last_stats(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        return {"iters": int(out[0]), "rr": out[1], "atol_sq": out[2], "scale": out[3]}

    def eval_terms(self, mask, q, qd, activations=None):
        P = self.P
        q = np.ascontiguousarray(q, dtype=np.float64).reshape(P, 3)
        qd = np.ascontiguousarray(qd, dtype=np.float64).reshape(P, 3)
        out = np.zeros((P, 3))
        act = None if activations is None else np.ascontiguousarray(activations, dtype=np.float64)
        dp = ctypes.POINTER(ctypes.c_double)
        lib().orc_eval_terms(self.h, int(mask), q.ctypes.data_as(dp), qd.ctypes.data_as(dp),
                             None if act is None else act.ctypes.data_as(ctypes.c_void_p), out.ctypes.data_as(dp))
        return out

    # ---- differentiation (fixed-matrix linearisation, differentiation.py) ----
    @staticmethod
    def _arr(a, shape=None):
        a = np.ascontiguousarray(a, dtype=np.float64)
        return a if shape is None else a.reshape(shape)

    def vjp_terms(self, mask, q, qd, weights, activations=None):
        """Reverse-mode product of eval_terms(mask) with per-node weights.

        Returns (dL/dq, dL/dqd, dL/dact [T], dL/d(mu, lambda, k_damp) [T, 3]) for
        L = sum_i weights_i . eval_terms(mask)_i.
        """
        P, T = self.P, self.T
        dp = ctypes.POINTER(ctypes.c_double)
        q, qd, w = self._arr(q, (P, 3)), self._arr(qd, (P, 3)), self._arr(weights, (P, 3))
        act = None if activations is None else self._arr(activations)
        xb, vb, ab, mb = np.zeros((P, 3)), np.zeros((P, 3)), np.zeros(max(T, 1)), np.zeros((max(T, 1), 3))
        lib().orc_vjp_terms(self.h, int(mask), q.ctypes.data_as(dp), qd.ctypes.data_as(dp),
                            None if act is None else act.ctypes.data_as(ctypes.c_void_p), w.ctypes.data_as(dp),
                            xb.ctypes.data_as(dp), vb.ctypes.data_as(dp), ab.ctypes.data_as(dp), mb.ctypes.data_as(dp))
        return xb, vb, ab[:T], mb[:T]

    def diff_step(self, q, qd, external_forces=None, activations=None, tol=1e-13):
        """Differentiable step with the current (fixed) matrix: returns (q_out, qd_out, dv)."""
        P = self.P
        dp = ctypes.POINTER(ctypes.c_double)
        q, qd = self._arr(q, (P, 3)), self._arr(qd, (P, 3))
        ext = None if external_forces is None else self._arr(external_forces, (P, 3))
        act = None if activations is None else self._arr(activations)
        qo, vo, dv = np.zeros((P, 3)), np.zeros((P, 3)), np.zeros((P, 3))
        lib().orc_diff_forward(self.h, q.ctypes.data_as(dp), qd.ctypes.data_as(dp),
                               None if ext is None else ext.ctypes.data_as(ctypes.c_void_p),
                               None if act is None else act.ctypes.data_as(ctypes.c_void_p), self.dt, float(tol),
                               qo.ctypes.data_as(dp), vo.ctypes.data_as(dp), dv.ctypes.data_as(dp))
        return qo, vo, dv

    def diff_step_backward(self, q, qd, dv, q_out_grad, qd_out_grad, activations=None, tol=1e-13):
        """Adjoint of diff_step: returns dict(q, qd, act, ext, mat) input gradients."""
        P, T = self.P, self.T
        dp = ctypes.POINTER(ctypes.c_double)
        q, qd, dv = self._arr(q, (P, 3)), self._arr(qd, (P, 3)), self._arr(dv, (P, 3))
        go, gv = self._arr(q_out_grad, (P, 3)), self._arr(qd_out_grad, (P, 3))
        act = None if activations is None else self._arr(activations)
        qb, vb, eb = np.zeros((P, 3)), np.zeros((P, 3)), np.zeros((P, 3))
        ab, mb = np.zeros(max(T, 1)), np.zeros((max(T, 1), 3))
        lib().orc_diff_backward(self.h, q.ctypes.data_as(dp), qd.ctypes.data_as(dp),
                                None if act is None else act.ctypes.data_as(ctypes.c_void_p), dv.ctypes.data_as(dp),
                                go.ctypes.data_as(dp), gv.ctypes.data_as(dp), self.dt, float(tol),
                                qb.ctypes.data_as(dp), vb.ctypes.data_as(dp), ab.ctypes.data_as(dp),
                                eb.ctypes.data_as(dp), mb.ctypes.data_as(dp))
        return {"q": qb, "qd": vb, "act": ab[:T], "ext": eb, "mat": mb[:T]}

    def set_tet_materials(self, materials):
        """Replace per-tet (mu, lambda, k_damp); re-assemble with refresh_tangent afterwards."""
        m = self._arr(materials, (self.T, 3))
        lib().orc_set_tet_materials(self.h, m.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
