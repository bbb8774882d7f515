"""Invariants of the reference's tests/test_solver.py on cuda:0 (ports of its
production-mechanics tests to this API):

- :203-215 a rigidly moved cube (Tet4 and Hex8) has zero elastic force, a
  symmetric positive definite assembled matrix and the rest volume 0.1^3;
- :372-379 Hex8 quadrature and material storage of the Hex8 cube;
- :381-422 the checked (differentiable, deferred checks) and the fast
  solver run the same physics on the same sparse matrix: equal states within
  the fast path's 1% residual, equal pattern and values, the fast path's true
  residual within its tolerance, checked residuals within 2e-5, and a block
  reserve bounded by the mesh connectivity.
"""

import numpy as np
import pytest
import torch

from helpers import cube
from test_gpu_parity import DEV, np_

pytestmark = pytest.mark.gpu

DT = 0.002


def make_solver(model, *, differentiable=True, dt=DT, **options):
    from newton_solver_fem_amd.solver import SolverFEM

    args = dict(dt=dt, differentiable=differentiable, linear_solver_maxiter=128, solver_type="cg",
                preconditioner_type="block_diag", self_contact_ke=0.0, min_stretch=-1.0)
    args.update(options)
    return SolverFEM(model, **args)


def elastic_force(solver, state):
    model = solver.model
    return np_(solver.eval_tetrahedral_forces(model, None, state)) + np_(
        solver.eval_hexahedral_forces(model, None, state))


@pytest.mark.parametrize("kind", ["tet", "hex"])
def test_rest_rigid_motion_and_hex_patch(kind):
    print(f"case invariants rigid {kind}", flush=True)
    rotation = np.array([[0.8, -0.6, 0.0], [0.6, 0.8, 0.0], [0.0, 0.0, 1.0]])
    model = cube(kind, device=DEV)
    solver = make_solver(model, differentiable=False)
    state = model.state()
    moved = model.host["particle_q"].astype(np.float64) @ rotation.T + (0.1, 0.2, -0.15)
    state.particle_q.copy_(torch.tensor(moved, dtype=torch.float32))
    np.testing.assert_allclose(elastic_force(solver, state), 0.0, atol=3e-4)
    solver.refresh_tangent(state.particle_q)
    matrix = solver.A_bsr.dense()
    np.testing.assert_allclose(matrix, matrix.T, atol=2e-8)
    assert np.linalg.eigvalsh(matrix).min() > 0.0
    assert abs(solver.compute_volume(state) - 0.1 ** 3) < 1e-8


def test_hex_quadrature_and_material_storage():
    print("case invariants hex storage", flush=True)
    model = cube("hex", device=DEV)
    solver = make_solver(model)
    assert solver.hex_count == 8
    assert tuple(model.fem.hex_materials.shape) == (8, 3)
    np.testing.assert_allclose(np_(model.fem.hex_materials), np.tile((500, 1000, 0), (8, 1)))
    assert abs(float(model.fem.hex_det_J0_w.double().sum()) - 0.1 ** 3) < 1e-8
    assert abs(float(model.particle_mass.double().sum()) - 100 * 0.1 ** 3) < 1e-6


def run(model, solver, q, v, forces):
    states = [model.state() for _ in range(len(forces) + 1)]
    states[0].particle_q.copy_(torch.tensor(q, dtype=torch.float32))
    states[0].particle_qd.copy_(torch.tensor(v, dtype=torch.float32))
    for i, f in enumerate(forces):
        solver.step(states[i], states[i + 1], None, None, solver.dt,
                    external_forces=torch.tensor(f, dtype=torch.float32, device=DEV))
    torch.cuda.synchronize()
    return states


def test_checked_and_fast_paths_use_same_physics_and_sparse_matrix():
    print("case invariants checked vs fast", flush=True)
    model = cube(device=DEV)
    checked = make_solver(model, defer_differentiation_checks=True)
    fast = make_solver(model, differentiable=False)
    q = model.host["particle_q"].astype(np.float64).copy()
    q[:, 0] *= 1.02
    velocity = np.tile((0.02, -0.03, 0.01), (len(q), 1)).astype(np.float32)
    forces = np.zeros((3, *q.shape), dtype=np.float32)
    forces[:, -1, 2] = -0.01
    for s in (checked, fast):
        s.refresh_tangent(torch.tensor(q, dtype=torch.float32, device=DEV))
    taped = run(model, checked, q, velocity, forces)
    ordinary = run(model, fast, q, velocity, forces)
    for a, c in zip(taped, ordinary):
        np.testing.assert_allclose(np_(a.particle_q), np_(c.particle_q), atol=2e-6)
        checked_v, fast_v = np_(a.particle_qd), np_(c.particle_qd)
        if np.linalg.norm(checked_v) > 0:
            # the fast solve stops at its 1% residual: compare whole-state error
            assert np.linalg.norm(checked_v - fast_v) / np.linalg.norm(checked_v) < 0.01
    matrix = fast.A_bsr.dense()
    for s_in, s_out in zip(ordinary[:-1], ordinary[1:]):
        delta = (np_(s_out.particle_qd) - np_(s_in.particle_qd)).ravel()
        rhs = np_(s_in.particle_f).ravel()
        assert np.linalg.norm(matrix @ delta - rhs) / np.linalg.norm(rhs) < 0.0101
    residuals = checked.check_differentiation()
    assert residuals["forward_relative_residual"] < 2e-5
    assert residuals["backward_relative_residual"] < 2e-5
    np.testing.assert_array_equal(np_(checked.A_bsr.offsets), np_(fast.A_bsr.offsets))
    active = int(checked.A_bsr.offsets[-1])
    np.testing.assert_array_equal(np_(checked.A_bsr.columns)[:active], np_(fast.A_bsr.columns)[:active])
    np.testing.assert_array_equal(np_(checked.A_bsr.values)[:active], np_(fast.A_bsr.values)[:active])
    assert active < model.particle_count ** 2
    reserve = model.particle_count + 16 * model.tet_count + 64 * model.fem.hex_count
    reserve += 3 * model.tri_count + 2 * model.spring_count
    assert checked.A_bsr.values.shape[0] <= reserve
