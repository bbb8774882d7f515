"""PreparedFEMStep on cuda:0 (reference tests/test_fem_prepared_step.py): the
prepared right-hand side, solve and reconstruction reproduce step(), the
operator / diagonal / preconditioner actions match the assembled matrix, the
right-hand side is affine in the load, preparation owns its input, and a
changed timestep invalidates the preparation."""

import numpy as np
import pytest
import torch

from newton_solver_fem_amd.model import FEMBuilder, Model
from test_gpu_parity import DEV, np_

pytestmark = pytest.mark.gpu
DT = 0.002


def fixture(mixed=True, fixed=True):
    b = FEMBuilder(gravity=(0.0, 0.0, -0.3))
    corners = np.array(((0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)))
    for p in 0.1 * corners:
        b.add_particle(p, (0, 0, 0), mass=0.001)
    b.add_tetrahedron(0, 1, 3, 4, k_mu=500, k_lambda=1000, k_damp=0.05)
    if mixed:
        b.add_soft_hex_mesh(np.arange(8).reshape(1, 8), k_mu=500, k_lambda=1000, k_damp=0.02)
        b.add_triangle(0, 1, 3, tri_ke=30, tri_ka=50, tri_kd=0.03)
        b.add_spring(2, 6, ke=40, kd=0.2, control=0)
    model = b.finalize(DEV)

    def solver():
        from newton_solver_fem_amd.solver import SolverFEM

        return SolverFEM(model, dt=DT, solver_type="cg", preconditioner_type="block_diag", linear_solver_maxiter=150,
                         linear_solver_tolerance=1e-7, self_contact_ke=0, linear_damping=0.4, min_stretch=0.98,
                         max_stretch=1.02, fixed_particle_indices=(0, 1) if fixed else None)

    current = model.state()
    q = np_(current.particle_q)
    q[4] += (0.003, -0.002, 0.009)
    q[6] += (-0.004, 0.003, -0.007)
    current.particle_q.copy_(torch.tensor(q, dtype=torch.float32))
    v = np.linspace(-0.02, 0.03, 24, dtype=np.float32).reshape(8, 3)
    if fixed:
        v[:2] = 0
    current.particle_qd.copy_(torch.tensor(v))
    current.particle_f.fill_(7.0)
    return model, current, solver


def test_mixed_elements_damping_clamps_load_and_fixed_support_parity():
    print("case prepared parity", flush=True)
    model, current, make = fixture()
    actual, reference = make(), make()
    ref_in, ref_out, result = model.state(), model.state(), model.state()
    for name in ("particle_q", "particle_qd", "particle_f"):
        getattr(ref_in, name).copy_(getattr(current, name))
    for s in (actual, reference):
        s.refresh_tangent(current.particle_q)
    before = [t.clone() for t in (current.particle_q, current.particle_qd, current.particle_f)]
    guess = torch.full((8, 3), 0.001, dtype=torch.float32, device=DEV)
    actual.velocity_increment.copy_(guess)
    reference.velocity_increment.copy_(guess)
    prepared = actual.prepare_step(current, None, None, DT)
    for scale in (0.0, 1.0, -0.5):
        ext = torch.tensor(scale * np.linspace(-0.01, 0.02, 24).reshape(8, 3), dtype=torch.float32, device=DEV)
        prepared.velocity_increment.copy_(guess)
        reference.velocity_increment.copy_(guess)
        rhs = prepared.rhs(ext)
        dv = prepared.solve(rhs)
        prepared.reconstruct(dv, result)
        reference.step(ref_in, ref_out, None, None, DT, external_forces=ext)
        np.testing.assert_allclose(np_(rhs), np_(ref_in.particle_f), rtol=2e-6, atol=2e-9)
        np.testing.assert_allclose(np_(dv), np_(reference.velocity_increment), rtol=3e-5, atol=2e-7)
        np.testing.assert_allclose(np_(result.particle_q), np_(ref_out.particle_q), rtol=1e-6, atol=1e-8)
        np.testing.assert_allclose(np_(result.particle_qd), np_(ref_out.particle_qd), rtol=3e-5, atol=2e-7)
        np.testing.assert_array_equal(np_(result.particle_q)[:2], model.host["particle_q"][:2])
        np.testing.assert_array_equal(np_(result.particle_qd)[:2], np.zeros((2, 3)))
    for t, b in zip((current.particle_q, current.particle_qd, current.particle_f), before):
        assert torch.equal(t, b)
    assert torch.equal(actual.velocity_increment, guess)


def test_operator_diagonal_rhs_affinity_and_reconstruction_without_solve():
    print("case prepared operator", flush=True)
    model, current, make = fixture(mixed=False)
    solver = make()
    solver.refresh_tangent(current.particle_q)
    prepared = solver.prepare_step(current, None, None, DT)
    matrix = solver.A_bsr.dense()
    fixed = prepared.fixed_mask.cpu().numpy().astype(bool)
    sf = np.repeat(fixed, 3)
    matrix[sf, :] = 0
    matrix[:, sf] = 0
    matrix[sf, sf] = 1
    vec_np = np.linspace(-0.001, 0.002, 24, dtype=np.float32).reshape(8, 3)
    vec = torch.tensor(vec_np, device=DEV)
    out = torch.empty_like(vec)
    prepared.apply_operator(vec, out)
    np.testing.assert_allclose(np_(out).ravel(), matrix @ vec_np.ravel().astype(np.float64), rtol=2e-6, atol=1e-9)
    diag = np.array([matrix[3 * i:3 * i + 3, 3 * i:3 * i + 3] for i in range(8)])
    np.testing.assert_allclose(np_(prepared.diagonal_blocks), diag, rtol=1e-7, atol=1e-9)
    prepared.apply_preconditioner(vec, out)
    minv = np_(solver._pc_inv_diag).reshape(8, 3, 3)
    np.testing.assert_allclose(np_(out), np.einsum("nij,nj->ni", minv, vec_np), rtol=1e-6, atol=1e-12)
    rhs = np_(prepared.rhs(vec))
    diff = DT * vec_np.astype(np.float64)
    diff[fixed] = 0
    np.testing.assert_allclose(rhs, np_(prepared.base_rhs) + diff, rtol=2e-6, atol=1e-9)
    result = model.state()
    prepared.reconstruct(vec, result)
    ev = np_(current.particle_qd) + vec_np
    eq = np_(current.particle_q) + DT * ev
    ev[fixed] = 0
    eq[fixed] = model.host["particle_q"][fixed]
    np.testing.assert_allclose(np_(result.particle_q), eq, rtol=1e-7, atol=1e-9)
    np.testing.assert_allclose(np_(result.particle_qd), ev, rtol=1e-7, atol=1e-9)
    with pytest.raises(ValueError, match="distinct"):
        prepared.apply_operator(vec, vec)


def test_preparation_owns_input_and_reuses_storage():
    print("case prepared ownership", flush=True)
    model, current, make = fixture(mixed=False)
    solver = make()
    prepared = solver.prepare_step(current, None, None, DT)
    pointers = (prepared.base_rhs.data_ptr(), prepared.velocity_increment.data_ptr(),
                prepared.diagonal_blocks.data_ptr())
    q, v = current.particle_q.clone(), current.particle_qd.clone()
    current.particle_q.fill_(3.0)
    current.particle_qd.fill_(4.0)
    zeros = torch.zeros_like(prepared.base_rhs)
    result = model.state()
    prepared.reconstruct(zeros, result)
    prepared.rhs(zeros)
    expected = np_(q) + DT * np_(v)
    expected[:2] = model.host["particle_q"][:2]
    np.testing.assert_allclose(np_(result.particle_q), expected, rtol=1e-7, atol=1e-9)
    current.particle_q.copy_(q)
    current.particle_qd.copy_(v)
    again = solver.prepare_step(current, None, None, DT)
    assert again is prepared
    assert pointers == (again.base_rhs.data_ptr(), again.velocity_increment.data_ptr(),
                        again.diagonal_blocks.data_ptr())


def test_changed_timestep_requires_refresh_and_invalidates_old_preparation():
    print("case prepared timestep", flush=True)
    model, current, make = fixture()
    solver = make()
    prepared = solver.prepare_step(current, None, None, DT)
    smaller = DT / 2
    with pytest.raises(ValueError, match="assembled tangent"):
        solver.prepare_step(current, None, None, smaller)
    solver.refresh_tangent(current.particle_q, smaller)
    with pytest.raises(RuntimeError, match="changed"):
        prepared.rhs()
    prepared = solver.prepare_step(current, None, None, smaller)
    result, reference = model.state(), model.state()
    dv = prepared.solve(prepared.base_rhs)
    prepared.reconstruct(dv, result)
    solver.step(current, reference, None, None, smaller)
    np.testing.assert_allclose(np_(result.particle_q), np_(reference.particle_q), rtol=1e-6, atol=1e-8)
    np.testing.assert_allclose(np_(result.particle_qd), np_(reference.particle_qd), rtol=3e-5, atol=2e-7)


def test_triangle_only_prepared_constant_load():
    print("case prepared triangle", flush=True)
    from newton_solver_fem_amd.solver import SolverFEM

    b = FEMBuilder(gravity=(0.0, 0.0, -0.1))
    for p in ((0, 0, 0), (0.1, 0, 0), (0, 0.1, 0)):
        b.add_particle(p, (0, 0, 0), mass=0.002)
    b.add_triangle(0, 1, 2, tri_ke=30, tri_ka=50, tri_kd=0.02)
    model = b.finalize(DEV)
    solver = SolverFEM(model, dt=DT, solver_type="cg", preconditioner_type="block_diag", self_contact_ke=0)
    current, actual, expected = model.state(), model.state(), model.state()
    current.particle_qd.copy_(torch.arange(9, dtype=torch.float32).reshape(3, 3) * 0.001)
    load = torch.full((3, 3), 0.002, dtype=torch.float32, device=DEV)
    prepared = solver.prepare_step(current, None, None, DT, external_forces=load)
    prepared.reconstruct(prepared.solve(prepared.base_rhs), actual)
    solver.step(current, expected, None, None, DT, external_forces=load)
    np.testing.assert_array_equal(np_(prepared.base_rhs), np_(current.particle_f))
    np.testing.assert_allclose(np_(actual.particle_qd), np_(expected.particle_qd), rtol=1e-6, atol=1e-8)


def test_prepared_actions_replay_in_a_hip_graph():
    """rhs + solve + reconstruct recorded once into a HIP graph and replayed
    (with eager work between replays) equal the eager sequence each time."""
    print("case prepared graph", flush=True)
    model, current, make = fixture()
    solver = make()
    solver.refresh_tangent(current.particle_q)
    prepared = solver.prepare_step(current, None, None, DT)
    ext = torch.tensor(np.linspace(-0.01, 0.02, 24).reshape(8, 3), dtype=torch.float32, device=DEV)
    rhs = torch.zeros_like(ext)
    guess = prepared.velocity_increment.clone()
    result = model.state()

    def body():
        prepared.rhs(ext, out=rhs)
        prepared.solve(rhs)
        prepared.reconstruct(prepared.velocity_increment, result)

    body()
    torch.cuda.synchronize()
    eager = (rhs.clone(), prepared.velocity_increment.clone(), result.particle_q.clone(), result.particle_qd.clone())
    prepared.velocity_increment.copy_(guess)
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        body()
    torch.cuda.current_stream().wait_stream(side)
    g = torch.cuda.CUDAGraph()
    prepared.velocity_increment.copy_(guess)
    with torch.cuda.graph(g):
        body()
    for k in range(3):
        prepared.velocity_increment.copy_(guess)
        rhs.zero_()
        result.particle_q.zero_()
        g.replay()
        torch.cuda.synchronize()
        got = (rhs, prepared.velocity_increment, result.particle_q, result.particle_qd)
        for a, b in zip(got, eager):
            assert torch.equal(a, b), k
        solver.step(current, model.state(), None, None, DT)  # eager work between replays
        torch.cuda.synchronize()


def test_interrupted_preparation_fails_closed_without_mutating_input():
    """Reference tests/test_fem_prepared_step.py:200-211: a preparation that
    fails part-way leaves the workspace unusable (not stale) and the caller's
    state untouched; preparing again recovers."""
    print("case prepared interrupted", flush=True)
    from unittest import mock

    _, current, make = fixture(mixed=False)
    solver = make()
    solver.refresh_tangent(current.particle_q)
    prepared = solver.prepare_step(current, None, None, DT)
    before = current.particle_q.clone()
    with mock.patch.object(type(prepared), "_call", side_effect=RuntimeError("injected failure")):
        with pytest.raises(RuntimeError, match="injected failure"):
            solver.prepare_step(current, None, None, DT)
    with pytest.raises(RuntimeError, match="Prepare FEM mechanics"):
        prepared.rhs()
    assert torch.equal(current.particle_q, before)
    solver.prepare_step(current, None, None, DT)
    assert torch.isfinite(prepared.rhs()).all()


def test_unsupported_modes_and_invalid_alias_fail_before_state_change():
    """Reference tests/test_fem_prepared_step.py:214-231 (the tape case has no
    counterpart here: differentiable solvers are rejected instead)."""
    print("case prepared unsupported", flush=True)
    from newton_solver_fem_amd.solver import SolverFEM

    model, current, make = fixture(mixed=False, fixed=False)
    diff = SolverFEM(model, dt=DT, solver_type="cg", preconditioner_type="block_diag", self_contact_ke=0,
                     differentiable=True)
    with pytest.raises(RuntimeError, match="non-differentiable"):
        diff.prepare_step(current, None, None, DT)
    solver = make()
    solver.refresh_tangent(current.particle_q)
    prepared = solver.prepare_step(current, None, None, DT)
    vector = torch.zeros_like(prepared.base_rhs)
    with pytest.raises(ValueError, match="alias"):
        prepared.rhs(vector, vector)
    with pytest.raises(ValueError, match="distinct"):
        prepared.apply_operator(vector, vector)
    solver = make()
    solver.prepare_tetrahedral_updates(model.particle_count)
    with pytest.raises(ValueError, match="assembled"):
        solver.prepare_step(current, None, None, DT)
