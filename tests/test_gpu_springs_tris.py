"""Springs, triangle membranes and mixed Tet4/Hex8/triangle/spring systems on
cuda:0 (through the C-ABI) vs the FP64 oracle: pattern bit-exact, constant and
element blocks, force terms (with damping, drag and lift), right-hand side
including the spring-damping predictor restore, and converged steps across
timestep changes (reference tests/test_solver_timestep.py:53-110).
"""

import numpy as np
import pytest
import torch

from newton_solver_fem_amd.model import FEMBuilder, ModelBuilder
from test_gpu_parity import DEV, np_, oracle_for, rel, solver_for
from test_oracle_springs_tris import mixed_model

pytestmark = pytest.mark.gpu


def on_device(model_cpu):
    return type(model_cpu)(DEV, model_cpu.host, model_cpu.gravity)


def test_mixed_timestep_change_matches_fresh_and_oracle():
    print("case mixed timestep", flush=True)
    model = on_device(mixed_model())
    kw = dict(linear_solver_maxiter=80, linear_solver_tolerance=1e-5, fixed_particle_indices=(0, 1))
    actual = solver_for(model, dt=0.002, **kw)
    ref_pattern = oracle_for(model, dt=0.002, fixed=(0, 1))
    off, cols, _, _ = ref_pattern.matrix()
    np.testing.assert_array_equal(actual.A_bsr.offsets.cpu().numpy(), off)
    np.testing.assert_array_equal(actual.A_bsr.columns.cpu().numpy(), cols)
    q = model.host["particle_q"].astype(np.float64).copy()
    q[2] += (0.002, -0.001, 0.001)
    q[6] += (-0.001, 0.002, -0.003)
    velocity = np.linspace(-0.02, 0.03, 24, dtype=np.float32).reshape(8, 3)
    velocity[:2] = 0
    force = torch.tensor(np.linspace(-0.01, 0.02, 24, dtype=np.float32).reshape(8, 3), device=DEV)
    current, following = model.state(), model.state()
    current.particle_q.copy_(torch.tensor(q, dtype=torch.float32))
    actual.refresh_tangent(current.particle_q)
    for dt in (0.002, 0.001, 0.002):
        fresh = solver_for(model, dt=dt, **kw)
        fin, fout = model.state(), model.state()
        for state, solver in ((current, actual), (fin, fresh)):
            state.particle_q.copy_(torch.tensor(q, dtype=torch.float32))
            state.particle_qd.copy_(torch.tensor(velocity))
            solver.velocity_increment.copy_(0.25 * torch.tensor(velocity, device=DEV))
            solver.refresh_tangent(state.particle_q, dt)
        assert actual.dt == dt
        np.testing.assert_allclose(actual.A_bsr.dense(), fresh.A_bsr.dense(), rtol=1e-6, atol=1e-9)
        ref = oracle_for(model, dt=dt, fixed=(0, 1), linear_solver_tolerance=1e-10, linear_solver_maxiter=200)
        ref.refresh_tangent(q, dt)
        _, _, vals, minv = ref.matrix()
        assert rel(np_(actual.A_bsr.values), vals) < 2e-6
        assert rel(np_(actual._pc_inv_diag), minv) < 2e-5
        fresh.velocity_increment.copy_(actual.velocity_increment)
        actual.step(current, following, None, None, dt, external_forces=force)
        fresh.step(fin, fout, None, None, dt, external_forces=force)
        np.testing.assert_allclose(np_(following.particle_qd), np_(fout.particle_qd), rtol=2e-5, atol=1e-7)
        np.testing.assert_allclose(np_(current.particle_f), np_(fin.particle_f), rtol=1e-6, atol=1e-9)
        out = ref.step(q, velocity.astype(np.float64), dt=dt, external_forces=np_(force))
        assert rel(np_(current.particle_f), out["f"]) < 1e-5
        np.testing.assert_allclose(np_(following.particle_qd), out["qd"], rtol=3e-4, atol=2e-6)
        np.testing.assert_allclose(np_(following.particle_q), out["q"], rtol=1e-6, atol=1e-8)


def cloth_model(device, drag=0.2, lift=0.5):
    """Soft Tet4 block (add_soft_mesh: tets, open-face triangles, unique-edge springs)."""
    b = FEMBuilder(gravity=(0.0, 0.0, -9.81))
    b.default_particle_radius = 0.001
    n = 3
    g = np.stack(np.meshgrid(*[np.arange(n + 1)] * 3, indexing="ij"), -1).reshape(-1, 3) * 0.02
    idx = lambda i, j, k: (i * (n + 1) + j) * (n + 1) + k
    tets = []
    for i in range(n):
        for j in range(n):
            for k in range(n):
                c = [idx(i + a, j + b_, k + d) for a, b_, d in
                     ((0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1))]
                for t in ((0, 1, 3, 4), (1, 2, 3, 6), (1, 4, 5, 6), (3, 4, 6, 7), (1, 3, 4, 6)):
                    tets.append([c[m] for m in t])
    b.add_soft_mesh(pos=(0, 0, 0.02), rot=(0, 0, 0, 1), scale=1.0, vel=(0.1, 0.0, -0.2), vertices=g,
                    indices=np.asarray(tets).ravel(), density=500.0, k_mu=2.0e3, k_lambda=8.0e3, k_damp=1.0,
                    tri_ke=50.0, tri_ka=80.0, tri_kd=0.5, tri_drag=drag, tri_lift=lift)
    return b.finalize(device)


def test_spring_and_triangle_forces_match_oracle():
    print("case spring/tri forces", flush=True)
    model = cloth_model(DEV)
    assert model.spring_count > 0 and model.tri_count > 0
    s = solver_for(model, dt=0.002)
    ref = oracle_for(model, dt=0.002)
    rng = np.random.default_rng(8)
    st = model.state()
    st.particle_q.add_(torch.tensor(rng.normal(0, 0.002, st.particle_q.shape), dtype=torch.float32, device=DEV))
    st.particle_qd.copy_(torch.tensor(rng.normal(0, 0.1, st.particle_qd.shape), dtype=torch.float32))
    q, v = np_(st.particle_q), np_(st.particle_qd)
    import oracle

    assert rel(np_(s.eval_spring_forces(model, st)), ref.eval_terms(oracle.TERM_SPRING, q, v)) < 1e-5
    assert rel(np_(s.eval_triangle_forces(model, None, st)), ref.eval_terms(oracle.TERM_TRI, q, v)) < 1e-5


def test_soft_mesh_drop_rollout():
    """Springs + triangles + tets falling onto the ground: 40 substeps, converged
    solves, positions within 1e-4 relative of the FP64 oracle."""
    print("case soft mesh rollout", flush=True)
    model = cloth_model(DEV)
    kw = dict(dt=1 / 600, linear_solver_tolerance=1e-7, linear_solver_maxiter=300, ground_plane=(0, 0, 1, 0),
              ground_ke=100.0, ground_kd=0.1, ground_kf=1.0, ground_mu=0.5)
    s = solver_for(model, **kw)
    ref = oracle_for(model, **kw)
    st0, st1 = model.state(), model.state()
    q, v = np_(st0.particle_q), np_(st0.particle_qd)
    for k in range(40):
        s.refresh_tangent(st0.particle_q)
        s.step(st0, st1, None, None, kw["dt"])
        ref.refresh_tangent(q)
        out = ref.step(q, v)
        q, v = out["q"], out["qd"]
        st0, st1 = st1, st0
    err = rel(np_(st0.particle_q), q)
    print(f"soft mesh rollout position rel err {err:.2e}, min z {q[:, 2].min():.4f}", flush=True)
    assert err < 1e-4
    assert q[:, 2].min() < 0.001  # the lowest nodes are in ground contact


def test_timestep_change_rules():
    """Reference test_solver_timestep.py:117-164 on the mixed model: a
    differentiable solver rejects a changed dt (matrix and dt unchanged), a
    changed dt clears only the warm start, and invalid dt values are rejected."""
    print("case timestep rules", flush=True)
    from newton_solver_fem_amd.solver import SolverFEM

    model = on_device(mixed_model())
    kw = dict(solver_type="cg", preconditioner_type="block_diag", linear_solver_maxiter=80,
              linear_solver_tolerance=1e-5, self_contact_ke=0.0)
    s = SolverFEM(model, dt=0.002, differentiable=True, **kw)
    before = s.A_bsr.dense()
    with pytest.raises(ValueError, match="Differentiable"):
        s.refresh_tangent(model.particle_q, 0.001)
    assert s.dt == 0.002
    np.testing.assert_array_equal(s.A_bsr.dense(), before)
    s = SolverFEM(model, dt=0.002, fixed_particle_indices=(0, 1), **kw)
    state = model.state()
    q0, v0 = state.particle_q.clone(), state.particle_qd.clone()
    guess = torch.full((8, 3), 0.02, dtype=torch.float32, device=DEV)
    s.velocity_increment.copy_(guess)
    s.refresh_tangent(state.particle_q)
    assert torch.equal(s.velocity_increment, guess)
    s.refresh_tangent(state.particle_q, 0.001)
    assert torch.equal(s.velocity_increment, torch.zeros_like(guess))
    assert torch.equal(state.particle_q, q0) and torch.equal(state.particle_qd, v0)
    before = s.A_bsr.dense()
    for dt in (0.0, -1.0, float("nan"), float("inf")):
        with pytest.raises(ValueError, match="finite and positive"):
            s.refresh_tangent(model.particle_q, dt)
        assert s.dt == 0.001
        np.testing.assert_array_equal(s.A_bsr.dense(), before)


def torque_chain():
    b = ModelBuilder(gravity=(0.0, 0.0, 0.0))
    rng = np.random.default_rng(21)
    pts = np.cumsum(rng.normal(0.0, 0.05, (8, 3)), axis=0)
    b.add_particles(pts, rng.normal(0.0, 0.2, (8, 3)), rng.uniform(0.01, 0.05, 8), radius=0.0)
    b.add_springs([(k, k + 1) for k in range(7)] + [(0, 4), (2, 6)], 30.0, 0.4)
    return b.finalize(DEV)


def test_spring_bend_torque_matches_oracle():
    """Bend torque (eval_springs_linear_and_torque kernels.py:222-292): spring forces and steps
    vs the oracle's linear spring force plus its torque restatement; short rest directions
    get no torque; a wrong shape is rejected like the reference (solver.py:396-400)."""
    import oracle

    print("case spring torque", flush=True)
    model = torque_chain()
    rng = np.random.default_rng(22)
    rest = rng.normal(size=(model.spring_count, 3))
    rest /= np.linalg.norm(rest, axis=1, keepdims=True)
    rest[3] *= 0.2  # torque-free spring
    kt, kw, dt = 2.5, 0.3, 1e-3
    s = solver_for(model, dt=dt, torque_stiffness=kt, torque_damping=kw, spring_rest_direction=rest,
                   linear_solver_tolerance=1e-7, linear_solver_maxiter=100)
    ref = oracle_for(model, dt=dt, linear_solver_tolerance=1e-12, linear_solver_maxiter=200)
    st = model.state()
    q, v = np_(st.particle_q), np_(st.particle_qd)
    idx = model.host["spring_indices"].reshape(-1, 2)
    f = np_(s.eval_spring_forces(model, st))
    expected = ref.eval_terms(oracle.TERM_SPRING, q, v) + oracle.spring_torque_forces(q, v, idx, rest, kt, kw)
    torque = oracle.spring_torque_forces(q, v, idx, rest, kt, kw)
    assert np.abs(torque).max() > 0.01 * np.abs(expected).max()
    assert rel(f, expected) < 1e-5
    a, b = model.state(), model.state()
    for k in range(3):
        s.refresh_tangent(a.particle_q)
        ref.refresh_tangent(q)
        out = ref.step(q, v, external_forces=oracle.spring_torque_forces(q, v, idx, rest, kt, kw))
        s.step(a, b, None, None, dt)
        assert rel(np_(b.particle_qd), out["qd"]) < 2e-5, k
        q, v = np_(b.particle_q), np_(b.particle_qd)
        a, b = b, a
    plain = solver_for(model, dt=dt, torque_stiffness=0.0, spring_rest_direction=rest)
    assert rel(np_(plain.eval_spring_forces(model, st)), ref.eval_terms(oracle.TERM_SPRING, np_(st.particle_q),
                                                                        np_(st.particle_qd))) < 1e-5
    with pytest.raises(ValueError, match="spring_rest_direction shape"):
        solver_for(model, dt=dt, torque_stiffness=kt, spring_rest_direction=rest[:-1])


def test_implicit_integration_equals_step_and_oracle():
    """SolverFEM.implicit_integration (reference solver.py:1254-1351) given
    state_in.particle_f = h (sum of the force terms + external load) gives the
    step's state, and the oracle's; the mixed model exercises the elastic
    predictor, the spring-damping restore and the fixed-row projection."""
    print("case implicit integration", flush=True)
    model = on_device(mixed_model())
    dt = 0.002
    kw = dict(linear_solver_maxiter=80, linear_solver_tolerance=1e-7, fixed_particle_indices=(0, 1))
    s = solver_for(model, dt=dt, **kw)
    ref = oracle_for(model, dt=dt, fixed=(0, 1), linear_solver_tolerance=1e-12, linear_solver_maxiter=200)
    q = model.host["particle_q"].astype(np.float64).copy()
    q[2] += (0.002, -0.001, 0.001)
    q[6] += (-0.001, 0.002, -0.003)
    v = np.linspace(-0.02, 0.03, 24).reshape(8, 3)
    v[:2] = 0
    ext = torch.tensor(np.linspace(-0.01, 0.02, 24, dtype=np.float32).reshape(8, 3), device=DEV)
    a, b, c = model.state(), model.state(), model.state()
    for st in (a, c):
        st.particle_q.copy_(torch.tensor(q, dtype=torch.float32))
        st.particle_qd.copy_(torch.tensor(v, dtype=torch.float32))
    s.refresh_tangent(a.particle_q)
    # h * (every force term + external load), then the second half of the step
    total = (s.eval_spring_forces(model, c) + s.eval_tetrahedral_forces(model, None, c)
             + s.eval_hexahedral_forces(model, None, c) + s.eval_triangle_forces(model, None, c)
             + s.eval_particle_ground_contact_forces(model, None, c) + s.eval_gravity_forces(model, c)
             + s.eval_linear_damping_forces(model, c) + ext)
    c.particle_f.copy_(total * dt)
    s.dv.zero_()
    s.implicit_integration(model, c, b, dt)
    torch.cuda.synchronize()
    q_ii, v_ii = np_(b.particle_q), np_(b.particle_qd)
    s.dv.zero_()
    s.step(a, b, None, None, dt, external_forces=ext)
    torch.cuda.synchronize()
    assert rel(v_ii, np_(b.particle_qd)) < 2e-5
    assert rel(q_ii, np_(b.particle_q)) < 1e-6
    ref.refresh_tangent(q)
    out = ref.step(q, v, external_forces=np_(ext))
    assert rel(v_ii, out["qd"]) < 2e-5
    assert rel(q_ii, out["q"]) < 1e-6
    np.testing.assert_array_equal(q_ii[:2], model.host["particle_q"][:2])  # fixed nodes at their anchors
    np.testing.assert_array_equal(v_ii[:2], 0.0)
