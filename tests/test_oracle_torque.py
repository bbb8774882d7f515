"""Pins the oracle's spring bend torque (oracle.spring_torque_forces, the torque
part of the reference's eval_springs_linear_and_torque kernels.py:222-292):

- with k_w = 0 the force on the first end is k_t L theta dtheta/dx_i, where
  theta = atan2(|d0 x d|, d0 . d) is computed independently here and its
  gradient by central differences (this fixes the factor and the sign the
  reference's formula gives), and the second end gets the opposite force;
- the damping term of a rigid rotation about the rotation axis is the closed
  form k_w (w . axis) (axis x d);
- no torque for a rest direction shorter than 0.5, an aligned spring or
  coincident ends.
"""

import numpy as np

import oracle


def theta_of(xi, xj, d0):
    d = (xi - xj) / np.linalg.norm(xi - xj)
    return np.arctan2(np.linalg.norm(np.cross(d0, d)), d0 @ d)


def test_stiffness_term_is_k_l_theta_times_angle_gradient():
    rng = np.random.default_rng(0)
    for _ in range(20):
        xi, xj = rng.normal(size=3), rng.normal(size=3)
        rest = rng.normal(size=3)
        rest /= np.linalg.norm(rest)
        q = np.stack([xi, xj])
        kt = 3.5
        f = oracle.spring_torque_forces(q, np.zeros_like(q), [(0, 1)], [rest], kt, 0.0)
        d0 = -rest
        eps = 1e-6
        g = np.zeros(3)
        for c in range(3):
            e = np.zeros(3)
            e[c] = eps
            g[c] = (theta_of(xi + e, xj, d0) - theta_of(xi - e, xj, d0)) / (2 * eps)
        L = np.linalg.norm(xi - xj)
        np.testing.assert_allclose(f[0], kt * L * theta_of(xi, xj, d0) * g, rtol=1e-6, atol=1e-8)
        np.testing.assert_allclose(f[1], -f[0], rtol=1e-12)


def test_damping_of_a_rigid_rotation():
    xi, xj = np.array([0.3, 0.1, 0.0]), np.array([-0.2, 0.0, 0.05])
    L = np.linalg.norm(xi - xj)
    d = (xi - xj) / L
    rest = -(d + np.array([0.0, 0.3, -0.2]))
    rest /= np.linalg.norm(rest)
    d0 = -rest
    axis = np.cross(d0, d)
    axis /= np.linalg.norm(axis)
    w = 2.0 * axis  # angular velocity about the rotation axis (perpendicular to d)
    v = np.stack([np.cross(w, xi - xj), np.zeros(3)])
    kw = 0.7
    f = oracle.spring_torque_forces(np.stack([xi, xj]), v, [(0, 1)], [rest], 0.0, kw)
    # tau = -kw (w . axis) L, F = tau (axis x d) / L, force on i = -F
    np.testing.assert_allclose(f[0], kw * (w @ axis) * np.cross(axis, d), rtol=1e-10)


def test_torque_free_cases():
    q = np.array([[1.0, 0.0, 0.0], [0.0, 0.0, 0.0]])
    v = np.zeros_like(q)
    # rest direction too short
    assert np.all(oracle.spring_torque_forces(q, v, [(0, 1)], [[0.0, 0.3, 0.0]], 5.0, 1.0) == 0)
    # aligned with the rest direction (-rest = d)
    assert np.all(oracle.spring_torque_forces(q, v, [(0, 1)], [[-1.0, 0.0, 0.0]], 5.0, 1.0) == 0)
    # coincident ends and ignored springs
    q2 = np.zeros((2, 3))
    assert np.all(oracle.spring_torque_forces(q2, v, [(0, 1)], [[0.0, 1.0, 0.0]], 5.0, 1.0) == 0)
    assert np.all(oracle.spring_torque_forces(q, v, [(-1, 1)], [[0.0, 1.0, 0.0]], 5.0, 1.0) == 0)
