"""Krylov methods besides PCG and the point-Jacobi preconditioners on cuda:0
(SolverFEM solver_type / preconditioner_type, reference solver.py:212-218,
:1142-1173, :1219-1252, which call warp.optim.linear), through the C-ABI:

- tiny-mass free fall for every method (reference
  tests/test_reference_block_preconditioner.py:71-91);
- live loads and warm starts equal the reduced dense solve for every method
  and preconditioner (reference tests/test_solver_cpu_batch.py:16-81);
- after a budget too small to converge, the iterate equals the oracle
  restatement's (oracle/krylov.py) on the same matrix, right-hand side,
  preconditioner and warm start;
- recorded steps replay changed loads and warm starts like eager steps
  (reference tests/test_solver_conditional_graph.py:99-130);
- the reference's default constructor (bicgstab + diag) on the Tet4 drop
  against the oracle's converged steps, and the true residual at tol 1e-2.
"""

import numpy as np
import pytest
import torch

import krylov
from helpers import cube
from newton_solver_fem_amd import ModelBuilder, build_tet_model, meshes
from test_gpu_parity import DEV, np_, oracle_for, rel, solver_for

pytestmark = pytest.mark.gpu

METHODS = ("cg", "bicgstab", "cr", "gmres")
PRECONDITIONERS = ("block_diag", "diag", "diag_abs", "id")


@pytest.mark.parametrize("method", METHODS)
def test_tiny_mass_falls_with_same_gravity(method):
    print(f"case krylov free fall {method}", flush=True)
    dt = 1 / 240
    for mass in (1e-25, 1e-18, 1e-14, 1.0):
        b = ModelBuilder(gravity=(0.0, 0.0, -9.81))
        b.add_particle((0, 0, 1), (0, 0, 0), mass=mass, radius=0.0)
        model = b.finalize(DEV)
        s = solver_for(model, dt=dt, solver_type=method, preconditioner_type="block_diag", linear_solver_maxiter=8)
        current, following = model.state(), model.state()
        for step in range(1, 4):
            s.step(current, following, model.control(), None, dt)
            np.testing.assert_allclose(np_(following.particle_qd)[0], (0, 0, -step * 9.81 * dt), atol=1e-7)
            np.testing.assert_allclose(np_(s.dv)[0], (0, 0, -9.81 * dt), atol=1e-7)
            np.testing.assert_allclose(np_(current.particle_f)[0], (0, 0, -mass * 9.81 * dt), rtol=3e-7, atol=0)
            current, following = following, current
        # A zero right-hand side clears the previous warm start.
        model.set_gravity((0, 0, 0))
        s.step(current, following, model.control(), None, dt)
        np.testing.assert_allclose(np_(following.particle_qd), np_(current.particle_qd), atol=1e-7)
        np.testing.assert_array_equal(np_(s.dv), np.zeros((1, 3)))


def four_particles(elastic=True):
    b = ModelBuilder(gravity=(0.0, 0.0, 0.0))
    for point in ((0, 0, 0), (0.125, 0, 0), (0, 0.125, 0), (0, 0, 0.125)):
        b.add_particle(point, (0.0, 0.0, 0.0), mass=0.001, radius=0.0)
    if elastic:
        b.add_tetrahedron(0, 1, 2, 3, k_mu=5000, k_lambda=10000, k_damp=0)
    return b.finalize(DEV)


LOADS = ((0, 2), (1, -3), (-2, 0.5), (0, 4))


def load(scale):
    return scale * np.array(((20, 30, 40), (10, -30, 60), (0.02, 0.04, 0.03), (-0.01, 0.03, 0.02)), np.float32)


@pytest.mark.parametrize("preconditioner", PRECONDITIONERS)
@pytest.mark.parametrize("method", METHODS)
def test_live_loads_and_warm_starts_match_reduced_solve(method, preconditioner):
    print(f"case krylov live loads {method} {preconditioner}", flush=True)
    model = four_particles()
    dt = 0.002
    s = solver_for(model, dt=dt, fixed_particle_indices=(0, 1), linear_solver_maxiter=80,
                   linear_solver_tolerance=1e-5, solver_type=method, preconditioner_type=preconditioner)
    matrix = s.A_bsr.dense()
    current, following = model.state(), model.state()
    elastic = np_(s.eval_tetrahedral_forces(model, model.control(), current))
    forces = torch.zeros(4, 3, dtype=torch.float32, device=DEV)
    for scale, warm_scale in LOADS:
        force = load(scale) - elastic.astype(np.float32)  # scale 0: an exactly zero right-hand side
        warm = np.full((4, 3), warm_scale * 0.01, dtype=np.float32)
        warm[:2] = 0
        forces.copy_(torch.tensor(force))
        s.velocity_increment.copy_(torch.tensor(warm))
        expected = np.zeros(12)
        expected[6:] = np.linalg.solve(matrix[6:, 6:], dt * (elastic + force.astype(np.float64)).ravel()[6:])
        s.step(current, following, None, None, dt, external_forces=forces)
        torch.cuda.synchronize()
        np.testing.assert_allclose(np_(s.velocity_increment).ravel(), expected, rtol=3e-4, atol=2e-6)
        np.testing.assert_array_equal(np_(following.particle_q)[:2], np_(model.particle_q)[:2])
        np.testing.assert_array_equal(np_(following.particle_qd)[:2], np.zeros((2, 3)))
        if scale == 0:
            assert s.solve_stats()["iterations"][0] == 0


def _preconditioner(s, A, kind):
    if kind == "block_diag":
        inv = np_(s._pc_inv_diag)
        M = np.zeros_like(A)
        for i in range(len(inv)):
            M[3 * i:3 * i + 3, 3 * i:3 * i + 3] = inv[i]
        return M
    return krylov.point_jacobi(A, kind)


@pytest.mark.parametrize("preconditioner", ("block_diag", "diag", "id"))
@pytest.mark.parametrize("method", ("bicgstab", "cr", "gmres"))
def test_unconverged_iterate_matches_oracle_restatement(method, preconditioner):
    print(f"case krylov iterate {method} {preconditioner}", flush=True)
    model = cube("tet", density=100.0, gravity=(0.0, 0.0, -9.81), device=DEV)
    fixed = (0, 1, 2)
    budget = 4
    s = solver_for(model, dt=0.002, fixed_particle_indices=fixed, linear_solver_maxiter=budget,
                   linear_solver_tolerance=1e-12, solver_type=method, preconditioner_type=preconditioner)
    rng = np.random.default_rng(11)
    st0, st1 = model.state(), model.state()
    st0.particle_q.add_(torch.tensor(rng.normal(0, 0.002, (27, 3)), dtype=torch.float32, device=DEV))
    st0.particle_qd.copy_(torch.tensor(rng.normal(0, 0.05, (27, 3)), dtype=torch.float32))
    s.refresh_tangent(st0.particle_q)
    warm = rng.normal(0, 0.01, (27, 3)).astype(np.float32)
    warm[list(fixed)] = 0
    s.velocity_increment.copy_(torch.tensor(warm))
    s.step(st0, st1, None, None, 0.002)
    torch.cuda.synchronize()
    A = s.A_bsr.dense()
    b = np_(st0.particle_f).ravel()
    M = _preconditioner(s, A, preconditioner)
    x, iters, _ = krylov.METHODS[method](krylov.projected_system(A, fixed), b, warm.astype(np.float64).ravel(), M,
                                         1e-12, budget)
    assert s.solve_stats()["iterations"][0] == iters
    assert rel(np_(s.velocity_increment).ravel(), x) < 2e-4


@pytest.mark.parametrize("method", ("cr", "bicgstab", "gmres"))
def test_recorded_steps_replay_changed_loads_and_warm_starts(method):
    print(f"case krylov graph {method}", flush=True)
    model = four_particles()
    dt = 0.002
    kw = dict(dt=dt, fixed_particle_indices=(0, 1), linear_solver_maxiter=80, linear_solver_tolerance=1e-5,
              solver_type=method, preconditioner_type="block_diag")
    actual, reference = solver_for(model, **kw), solver_for(model, **kw)
    current, following = model.state(), model.state()
    ref_in, ref_out = model.state(), model.state()
    external = torch.zeros(4, 3, dtype=torch.float32, device=DEV)
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        actual.step(current, following, None, None, dt, external_forces=external)
    torch.cuda.current_stream().wait_stream(side)
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        actual.step(current, following, None, None, dt, external_forces=external)
    for scale, warm_scale in ((0.0, 0.0), (1.0, 0.0), (-2.0, 0.5), (0.1, -3.0), (0.0, 1.0)):
        warm = np.full((4, 3), warm_scale * 0.01, dtype=np.float32)
        warm[:2] = 0
        external.copy_(torch.tensor(load(scale)))
        actual.velocity_increment.copy_(torch.tensor(warm))
        reference.velocity_increment.copy_(torch.tensor(warm))
        graph.replay()
        reference.step(ref_in, ref_out, None, None, dt, external_forces=external)
        torch.cuda.synchronize()
        np.testing.assert_allclose(np_(actual.velocity_increment), np_(reference.velocity_increment), atol=2e-8,
                                   rtol=2e-6)
        np.testing.assert_allclose(np_(following.particle_qd), np_(ref_out.particle_qd), atol=2e-8, rtol=2e-6)


@pytest.mark.parametrize("method", METHODS)
def test_default_constructor_drop_matches_oracle(method):
    """SolverFEM(model) with the reference's defaults except the method under
    test (preconditioner 'diag'): converged substeps of the Tet4 drop equal
    the oracle's; at tol 1e-2 the true relative residual meets the tolerance."""
    print(f"case krylov drop {method}", flush=True)
    from newton_solver_fem_amd.solver import SolverFEM

    scene = meshes.tet_drop_scene(n=6, height=0.002)
    model = build_tet_model(scene, device=DEV)
    ground = dict(ground_plane=(0.0, 0.0, 1.0, 0.0), ground_ke=100.0, ground_kd=0.1, ground_kf=1.0, ground_mu=0.5)
    dt = 1 / 600
    s = SolverFEM(model, dt=dt, solver_type=method, linear_solver_maxiter=300, linear_solver_tolerance=1e-6,
                  self_contact_ke=0.0, **ground)
    assert s.preconditioner_type == "diag"
    ref = oracle_for(model, dt=dt, linear_solver_tolerance=1e-11, linear_solver_maxiter=600, **ground)
    a, b = model.state(), model.state()
    a.particle_qd[:, 2] = -0.3
    q, v = np_(a.particle_q), np_(a.particle_qd)
    for _ in range(12):
        s.refresh_tangent(a.particle_q)
        s.step(a, b, None, None, dt)
        ref.refresh_tangent(q)
        out = ref.step(q, v)
        q, v = out["q"], out["qd"]
        a, b = b, a
    assert rel(np_(a.particle_qd), v) < 1e-4
    assert rel(np_(a.particle_q), q) < 1e-5
    # True residual at the default tolerance (reference tests/test_solver.py:403-405).
    s.linear_solver_tolerance = 1e-2
    s.refresh_tangent(a.particle_q)
    s.velocity_increment.zero_()
    s.step(a, b, None, None, dt)
    torch.cuda.synchronize()
    A = s.A_bsr.dense()
    rhs = np_(a.particle_f).ravel()
    r = A @ np_(s.velocity_increment).ravel() - rhs
    assert np.linalg.norm(r) / np.linalg.norm(rhs) < 0.0101
