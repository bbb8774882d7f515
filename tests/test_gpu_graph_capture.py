"""Captured substeps with the device-side residual exit replay changed loads,
warm starts and tangents exactly like an eager fixed-budget solve
(reference tests/test_solver_conditional_graph.py:99-180, CG).

The captured solver stops at the tolerance inside the graph
(linear_solver_use_cuda_graph=True); the eager reference runs the whole
budget with residual-masked updates (linear_solver_use_cuda_graph=False).
Both must give the same increment (rtol 2e-6 / atol 2e-8, the reference's
tolerances), the same right-hand side bit for bit and exact anchors.
"""

import numpy as np
import pytest
import torch

from newton_solver_fem_amd import ModelBuilder
from test_gpu_parity import DEV, solver_for

pytestmark = pytest.mark.gpu

DT = 0.002


def model(elastic=False):
    b = ModelBuilder(gravity=(0.0, 0.0, 0.0))
    for point in ((0, 0, 0), (0.1, 0, 0), (0, 0.1, 0), (0, 0, 0.1)):
        b.add_particle(point, (0.0, 0.0, 0.0), mass=0.001, radius=0.0)
    if elastic:
        b.add_tetrahedron(0, 1, 2, 3, k_mu=5000, k_lambda=10000, k_damp=0)
    return b.finalize(DEV)


def solver(m, **kw):
    args = dict(dt=DT, fixed_particle_indices=(0, 1), linear_solver_maxiter=80, linear_solver_tolerance=1e-5)
    args.update(kw)
    return solver_for(m, **args)


def capture(fn):
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        fn()
    torch.cuda.current_stream().wait_stream(side)
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        fn()
    return g


def t(a):
    return torch.tensor(np.asarray(a, np.float32), device=DEV)


def test_zero_first_capture_replays_changed_loads_and_warm_starts():
    print("case capture zero-first", flush=True)
    m = model()
    actual, reference = solver(m), solver(m, linear_solver_use_cuda_graph=False)
    current, following = m.state(), m.state()
    reference_in, reference_out = m.state(), m.state()
    external = torch.zeros(4, 3, dtype=torch.float32, device=DEV)
    actual.step(current, following, None, None, DT, external_forces=external)
    assert actual.solve_stats()["iterations"][0] == 0
    g = capture(lambda: actual.step(current, following, None, None, DT, external_forces=external))
    for scale, warm_scale in ((0.0, 0.0), (1.0, 0.0), (-2.0, 0.5), (0.1, -3.0), (0.0, 1.0)):
        force = scale * np.array(((20, 30, 40), (10, -30, 60), (0.02, 0.04, 0.03), (-0.01, 0.03, 0.02)), np.float32)
        warm = warm_scale * np.full((4, 3), 0.01, dtype=np.float32)
        warm[:2] = 0  # a constrained update never has fixed-node increments
        external.copy_(t(force))
        actual.velocity_increment.copy_(t(warm))
        reference.velocity_increment.copy_(t(warm))
        g.replay()
        stats = actual.solve_stats()
        reference.step(reference_in, reference_out, None, None, DT, external_forces=external)
        torch.cuda.synchronize()
        np.testing.assert_allclose(actual.velocity_increment.cpu().numpy(), reference.velocity_increment.cpu().numpy(),
                                   atol=2e-8, rtol=2e-6)
        np.testing.assert_allclose(following.particle_qd.cpu().numpy(), reference_out.particle_qd.cpu().numpy(),
                                   atol=2e-8, rtol=2e-6)
        np.testing.assert_array_equal(current.particle_f.cpu().numpy(), reference_in.particle_f.cpu().numpy())
        np.testing.assert_array_equal(following.particle_q.cpu().numpy()[:2], m.particle_q.cpu().numpy()[:2])
        np.testing.assert_array_equal(following.particle_qd.cpu().numpy()[:2], np.zeros((2, 3)))
        iterations = stats["iterations"][0]
        assert iterations < actual.linear_solver_maxiter
        assert stats["residual_sq"][0] <= stats["atol_sq"][0]
        assert (iterations > 0) if scale else (iterations == 0)


def test_captured_elastic_update_matches_fixed_budget_with_live_tangent():
    print("case capture elastic", flush=True)
    m = model(elastic=True)
    actual, reference = solver(m), solver(m, linear_solver_use_cuda_graph=False)
    current, following = m.state(), m.state()
    reference_in, reference_out = m.state(), m.state()
    external = torch.zeros(4, 3, dtype=torch.float32, device=DEV)
    actual.refresh_tangent(current.particle_q)
    actual.step(current, following, None, None, DT, external_forces=external)

    def body():
        actual.refresh_tangent(current.particle_q)
        actual.step(current, following, None, None, DT, external_forces=external)

    g = capture(body)
    rest = m.particle_q.cpu().numpy().copy()
    for scale in (1.0, -0.5, 2.0):
        q = rest.copy()
        q[2] += scale * np.array((0.002, 0.001, -0.001), np.float32)
        velocity = scale * np.array(((0, 0, 0), (0, 0, 0), (0.02, -0.03, 0.01), (-0.01, 0.01, 0.02)), np.float32)
        force = scale * np.array(((20, 30, 40), (10, -30, 60), (0.02, 0.04, 0.03), (-0.01, 0.03, 0.02)), np.float32)
        external.copy_(t(force))
        for state, s in ((current, actual), (reference_in, reference)):
            state.particle_q.copy_(t(q))
            state.particle_qd.copy_(t(velocity))
            s.velocity_increment.copy_(t(0.25 * velocity))
        g.replay()
        stats = actual.solve_stats()
        reference.refresh_tangent(reference_in.particle_q)
        reference.step(reference_in, reference_out, None, None, DT, external_forces=external)
        torch.cuda.synchronize()
        np.testing.assert_allclose(actual.velocity_increment.cpu().numpy(), reference.velocity_increment.cpu().numpy(),
                                   rtol=2e-6, atol=2e-8)
        np.testing.assert_allclose(following.particle_q.cpu().numpy(), reference_out.particle_q.cpu().numpy(),
                                   atol=2e-8, rtol=2e-6)
        np.testing.assert_allclose(following.particle_qd.cpu().numpy(), reference_out.particle_qd.cpu().numpy(),
                                   atol=2e-8, rtol=2e-6)
        np.testing.assert_array_equal(current.particle_f.cpu().numpy(), reference_in.particle_f.cpu().numpy())
        np.testing.assert_array_equal(following.particle_q.cpu().numpy()[:2], rest[:2])
        np.testing.assert_array_equal(following.particle_qd.cpu().numpy()[:2], np.zeros((2, 3)))
        assert stats["iterations"][0] < actual.linear_solver_maxiter
        assert stats["residual_sq"][0] <= stats["atol_sq"][0]


def particle_solver(mass, dt):
    b = ModelBuilder(gravity=(0.0, 0.0, -9.81))
    b.add_particle((0.0, 0.0, 1.0), (0.0, 0.0, 0.0), mass=mass, radius=0.0)
    m = b.finalize(DEV)
    return m, solver_for(m, dt=dt, linear_solver_maxiter=8)


@pytest.mark.parametrize("mass", [1e-25, 1e-18, 1e-14, 1.0])
def test_tiny_physical_mass_falls_with_same_gravity(mass):
    """Reference tests/test_reference_block_preconditioner.py:71-91: the
    right-hand-side rescaling keeps free fall exact down to 1e-25 kg, and a
    zero right-hand side clears the warm start."""
    print(f"case tiny mass {mass}", flush=True)
    dt = 1 / 240
    m, s = particle_solver(mass, dt)
    current, following = m.state(), m.state()
    pointers = (s.dv.data_ptr(), s._pc_inv_diag.data_ptr())
    for step in range(1, 4):
        s.step(current, following, m.control(), None, dt)
        np.testing.assert_allclose(following.particle_qd.cpu().numpy()[0], (0, 0, -step * 9.81 * dt), atol=1e-7)
        np.testing.assert_allclose(s.dv.cpu().numpy()[0], (0, 0, -9.81 * dt), atol=1e-7)
        np.testing.assert_allclose(current.particle_f.cpu().numpy()[0], (0, 0, -mass * 9.81 * dt), rtol=3e-7, atol=0)
        assert pointers == (s.dv.data_ptr(), s._pc_inv_diag.data_ptr())
        current, following = following, current
    m.set_gravity((0, 0, 0))
    s.step(current, following, m.control(), None, dt)
    np.testing.assert_allclose(following.particle_qd.cpu().numpy(), current.particle_qd.cpu().numpy(), atol=1e-7)
    np.testing.assert_array_equal(s.dv.cpu().numpy(), np.zeros((1, 3)))


def test_captured_tiny_rhs_replays_with_live_gravity_and_stable_buffers():
    """Reference tests/test_reference_block_preconditioner.py:93-108: gravity
    changed between replays of a recorded step is read live by the graph."""
    print("case capture live gravity", flush=True)
    dt = 1 / 240
    m, s = particle_solver(1e-25, dt)
    current, following = m.state(), m.state()
    s.step(current, following, m.control(), None, dt)
    pointers = (s.dv.data_ptr(), s._pc_inv_diag.data_ptr(), m.gravity_array.data_ptr())
    g = capture(lambda: s.step(current, following, None, None, dt))
    for gravity in (-9.81, 0.0, -4.0):
        m.set_gravity((0, 0, gravity))
        g.replay()
        torch.cuda.synchronize()
        np.testing.assert_allclose(following.particle_qd.cpu().numpy()[0], (0, 0, gravity * dt), atol=1e-7)
        np.testing.assert_allclose(s.dv.cpu().numpy()[0], (0, 0, gravity * dt), atol=1e-7)
        assert pointers == (s.dv.data_ptr(), s._pc_inv_diag.data_ptr(), m.gravity_array.data_ptr())
