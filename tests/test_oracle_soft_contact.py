"""Pins the oracle's particle-shape soft contacts (reference eval_soft_contacts
kernels.py:1199-1252) without the reference, which has no test of them:

- the penalty is -grad of E = 1/2 ke sum_k max(pen_k, 0)^2 over positions
  (central differences of an independently written energy);
- damping acts only against approach, along the normal, with kd |vn|;
- Coulomb friction: |ft| = mu fn when sliding, kf |vt| when sticking, opposite
  to the tangential relative velocity; none below the 1e-6 speed threshold;
- invariance under a rigid rotation of the whole scene and under a common
  velocity of particle and body (only the relative velocity counts);
- superposition over a particle's contacts; entries past the count, with a
  negative particle or without penetration contribute nothing.
"""

import numpy as np
import pytest

import oracle

KE, KD, KF, MU = 900.0, 12.0, 700.0, 0.4


def _scene(seed=0, n_particles=12, n_contacts=40):
    rng = np.random.default_rng(seed)
    q = rng.normal(0, 1.0, (n_particles, 3))
    qd = rng.normal(0, 0.5, (n_particles, 3))
    r = rng.uniform(0.1, 0.3, n_particles)
    particle = rng.integers(0, n_particles, n_contacts).astype(np.int32)
    normal = rng.normal(0, 1, (n_contacts, 3))
    normal /= np.linalg.norm(normal, axis=1, keepdims=True)
    # surface points slightly inside or outside the particle's radius
    depth = rng.uniform(-0.05, 0.15, n_contacts)
    body_pos = q[particle] - normal * (r[particle] - depth)[:, None]
    body_vel = rng.normal(0, 0.3, (n_contacts, 3))
    return q, qd, r, particle, body_pos, body_vel, normal


def _forces(q, qd, r, particle, bp, bv, nrm, count=None, ke=KE, kd=KD, kf=KF, mu=MU):
    count = len(particle) if count is None else count
    return oracle.soft_contact_forces(q, qd, r, count, particle, bp, bv, nrm, ke, kd, kf, mu)


def _energy(q, r, particle, bp, nrm, ke):
    e = 0.0
    for k in range(len(particle)):
        p = particle[k]
        gap = r[p] - (q[p] - bp[k]) @ nrm[k]
        e += 0.5 * ke * max(gap, 0.0) ** 2
    return e


def test_penalty_is_minus_energy_gradient():
    q, _, r, particle, bp, _, nrm = _scene(1)
    zero = np.zeros_like(q)
    f = _forces(q, zero, r, particle, bp, np.zeros_like(bp), nrm, kd=0.0)
    assert np.abs(f).max() > 1.0
    eps = 1e-6
    fd = np.zeros_like(q)
    for i in range(len(q)):
        for c in range(3):
            qp, qm = q.copy(), q.copy()
            qp[i, c] += eps
            qm[i, c] -= eps
            fd[i, c] = -(_energy(qp, r, particle, bp, nrm, KE) - _energy(qm, r, particle, bp, nrm, KE)) / (2 * eps)
    np.testing.assert_allclose(f, fd, rtol=1e-6, atol=1e-5)


def _one(pen, v, body_v=(0.0, 0.0, 0.0), **kw):
    # one particle of radius 0.2 touching the plane z = 0 (normal +z) with penetration pen
    q = np.array([[0.0, 0.0, 0.2 - pen]])
    return _forces(q, np.array([v], float), np.array([0.2]), np.array([0], np.int32), np.zeros((1, 3)),
                   np.array([body_v], float), np.array([[0.0, 0.0, 1.0]]), **kw)[0]


def test_damping_only_against_approach():
    pen = 0.01
    np.testing.assert_allclose(_one(pen, (0, 0, 0)), (0, 0, KE * pen), rtol=1e-14)
    np.testing.assert_allclose(_one(pen, (0, 0, -2.0)), (0, 0, KE * pen + KD * 2.0), rtol=1e-14)
    np.testing.assert_allclose(_one(pen, (0, 0, 3.0)), (0, 0, KE * pen), rtol=1e-14)  # separating: no pull


def test_coulomb_friction_regimes():
    pen = 0.02
    fn = KE * pen
    # sliding: kf |vt| > mu fn -> |ft| = mu fn, opposite to vt
    f = _one(pen, (1.0, 0.0, 0.0))
    np.testing.assert_allclose(f, (-MU * fn, 0.0, fn), rtol=1e-14)
    # sticking: tiny tangential speed -> |ft| = kf |vt|
    vt = 1e-3
    f = _one(pen, (0.0, vt, 0.0))
    assert KF * vt < MU * fn
    np.testing.assert_allclose(f, (0.0, -KF * vt, fn), rtol=1e-12)
    # below the speed threshold: normal force only
    np.testing.assert_allclose(_one(pen, (5e-7, 0.0, 0.0)), (0.0, 0.0, fn), rtol=1e-14)
    # friction never exceeds mu times the (nonnegative) normal force
    q, qd, r, particle, bp, bv, nrm = _scene(2, n_particles=30, n_contacts=30)
    particle = np.arange(30, dtype=np.int32)
    bp = q - nrm * (r - 0.05)[:, None]
    f = _forces(q, qd, r, particle, bp, bv, nrm)
    fn = np.sum(f * nrm, axis=1)
    ft = np.linalg.norm(f - fn[:, None] * nrm, axis=1)
    assert np.all(ft <= MU * np.maximum(fn, 0.0) * (1 + 1e-12) + 1e-12)


def test_only_relative_velocity_counts():
    q, qd, r, particle, bp, bv, nrm = _scene(3)
    f = _forces(q, qd, r, particle, bp, bv, nrm)
    shift = np.array([0.7, -1.1, 0.4])
    g = _forces(q, qd + shift, r, particle, bp, bv + shift, nrm)
    np.testing.assert_allclose(g, f, rtol=1e-12, atol=1e-10)


def test_rotation_invariance():
    q, qd, r, particle, bp, bv, nrm = _scene(4)
    f = _forces(q, qd, r, particle, bp, bv, nrm)
    rng = np.random.default_rng(9)
    R, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    g = _forces(q @ R.T, qd @ R.T, r, particle, bp @ R.T, bv @ R.T, nrm @ R.T)
    np.testing.assert_allclose(g, f @ R.T, rtol=1e-10, atol=1e-9)


def test_superposition_and_skips():
    q, qd, r, particle, bp, bv, nrm = _scene(5)
    f = _forces(q, qd, r, particle, bp, bv, nrm)
    g = np.zeros_like(f)
    for k in range(len(particle)):
        g += _forces(q, qd, r, particle[k:k + 1], bp[k:k + 1], bv[k:k + 1], nrm[k:k + 1])
    np.testing.assert_allclose(g, f, rtol=1e-12, atol=1e-10)
    # entries past the count and negative particles are skipped
    cut = 25
    h = _forces(q, qd, r, particle, bp, bv, nrm, count=cut)
    np.testing.assert_allclose(h, _forces(q, qd, r, particle[:cut], bp[:cut], bv[:cut], nrm[:cut]), rtol=0, atol=0)
    neg = particle.copy()
    neg[::2] = -1
    h = _forces(q, qd, r, neg, bp, bv, nrm)
    keep = neg >= 0
    np.testing.assert_allclose(h, _forces(q, qd, r, particle[keep], bp[keep], bv[keep], nrm[keep]), rtol=0, atol=0)
    # no penetration, no force
    far = bp - nrm * 10.0
    assert not np.any(_forces(q, qd, r, particle, far, bv, nrm))


@pytest.mark.parametrize("seed", [6, 7])
def test_order_of_contacts_only_rounds(seed):
    q, qd, r, particle, bp, bv, nrm = _scene(seed)
    perm = np.random.default_rng(seed).permutation(len(particle))
    f = _forces(q, qd, r, particle, bp, bv, nrm)
    g = _forces(q, qd, r, particle[perm], bp[perm], bv[perm], nrm[perm])
    np.testing.assert_allclose(g, f, rtol=1e-12, atol=1e-10)
