"""Pins the oracle's particle-particle contact (reference eval_particle_forces
kernels.py:1069-1114, particle_force :1043-1066) without the reference:

- the penalty term is -grad of E = 1/2 ke sum_pairs gap^2 (central
  differences of an independently written energy), with and without a
  cohesion distance;
- damping and Coulomb friction against closed forms of two-particle cases;
- pair forces are antisymmetric (the sum over particles is zero);
- inactive particles and co-located pairs contribute nothing.
"""

import numpy as np
import pytest

import oracle


def _cloud(n=40, seed=0, spacing=1.0):
    rng = np.random.default_rng(seed)
    q = rng.uniform(0, spacing * n ** (1 / 3), (n, 3))
    r = rng.uniform(0.3, 0.6, n) * spacing
    return q, r


def _pair_energy(q, r, ke, cohesion):
    # E = 1/2 ke sum_{i<j, gap <= cohesion} gap^2 with gap = |x_i - x_j| - r_i - r_j
    e = 0.0
    n = len(q)
    for i in range(n):
        for j in range(i + 1, n):
            gap = np.linalg.norm(q[i] - q[j]) - r[i] - r[j]
            if gap <= cohesion:
                e += 0.5 * ke * gap * gap
    return e


@pytest.mark.parametrize("cohesion", [0.0, 0.05])
def test_penalty_is_minus_energy_gradient(cohesion):
    q, r = _cloud()
    flags = np.ones(len(q), np.int32)
    ke = 7.0
    f = oracle.particle_contact_forces(q, np.zeros_like(q), r, flags, ke, 0.0, 0.0, 0.5, cohesion)
    assert np.abs(f).max() > 0.1  # some pairs overlap
    eps = 1e-6
    fd = np.zeros_like(q)
    for i in range(len(q)):
        for c in range(3):
            qp, qm = q.copy(), q.copy()
            qp[i, c] += eps
            qm[i, c] -= eps
            fd[i, c] = -(_pair_energy(qp, r, ke, cohesion) - _pair_energy(qm, r, ke, cohesion)) / (2 * eps)
    # Pairs exactly at the cohesion cut-off would make E jump; none is within eps of it here.
    np.testing.assert_allclose(f, fd, rtol=1e-6, atol=1e-6)


def test_head_on_damping_closed_form():
    # Two overlapping particles on the x axis closing at speed 2: fn = gap ke + vn kd with vn = -2.
    q = np.array([[0.0, 0.0, 0.0], [0.9, 0.0, 0.0]])
    v = np.array([[1.0, 0.0, 0.0], [-1.0, 0.0, 0.0]])
    r = np.array([0.5, 0.5])
    ke, kd = 100.0, 3.0
    f = oracle.particle_contact_forces(q, v, r, np.ones(2, np.int32), ke, kd, 0.0, 0.5)
    gap = 0.9 - 1.0
    fn = gap * ke + (-2.0) * kd  # n = -x for particle 0, n.v = -2
    np.testing.assert_allclose(f[0], [fn, 0.0, 0.0], rtol=1e-12)  # -n fn = +x fn
    np.testing.assert_allclose(f[1], -f[0], rtol=1e-12)
    # separating pair: no damping (min(vn, 0) = 0)
    f = oracle.particle_contact_forces(q, -v, r, np.ones(2, np.int32), ke, kd, 0.0, 0.5)
    np.testing.assert_allclose(f[0], [gap * ke, 0.0, 0.0], rtol=1e-12)


def test_coulomb_friction_closed_form():
    # Overlapping pair sliding tangentially (y): ft = min(|vt| kf, mu |fn|) along -vt.
    q = np.array([[0.0, 0.0, 0.0], [0.0, 0.0, 0.95]])
    r = np.array([0.5, 0.5])
    ke, mu = 200.0, 0.3
    fn = -0.05 * ke
    for kf, vy in ((1.0, 0.5), (1.0e4, 0.5)):  # below and at the Coulomb cap
        v = np.array([[0.0, vy, 0.0], [0.0, 0.0, 0.0]])
        f = oracle.particle_contact_forces(q, v, r, np.ones(2, np.int32), ke, 0.0, kf, mu)
        ft = min(vy * kf, mu * abs(fn))
        # n points from particle 1 to particle 0: -z
        np.testing.assert_allclose(f[0], [0.0, -ft, fn], rtol=1e-12, atol=1e-15)


def test_pair_forces_cancel_and_inactive_particles_are_skipped():
    q, r = _cloud(60, seed=3)
    rng = np.random.default_rng(5)
    v = rng.normal(size=q.shape)
    flags = np.ones(len(q), np.int32)
    f = oracle.particle_contact_forces(q, v, r, flags, 50.0, 2.0, 10.0, 0.4, 0.02)
    np.testing.assert_allclose(f.sum(0), 0.0, atol=1e-10 * np.abs(f).max())
    flags[::3] = 0
    g = oracle.particle_contact_forces(q, v, r, flags, 50.0, 2.0, 10.0, 0.4, 0.02)
    assert np.all(g[::3] == 0.0)
    sub = np.nonzero(flags)[0]
    h = oracle.particle_contact_forces(q[sub], v[sub], r[sub], flags[sub], 50.0, 2.0, 10.0, 0.4, 0.02)
    np.testing.assert_allclose(g[sub], h, rtol=1e-12, atol=1e-12)
    # co-located particles have no contact direction
    q2 = np.zeros((2, 3))
    f2 = oracle.particle_contact_forces(q2, np.zeros((2, 3)), np.ones(2), np.ones(2, np.int32), 1.0, 1.0, 1.0, 1.0)
    assert np.all(f2 == 0.0)
