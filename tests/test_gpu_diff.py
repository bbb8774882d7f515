"""Differentiable substep on the HIP path vs the FP64 oracle (fixed-matrix linearisation).

Forward: one differentiable step equals the oracle's (checked solve to 1e-8;
positions 1e-6, velocities 1e-4 relative).  Backward: every input gradient
(positions, velocities, activations, external forces, Tet4 materials) of a
step equals the oracle's reverse product within 1e-3 relative (fp32 vs fp64).
Rollouts: directional derivatives of a 12-step loss agree with FP64 central
differences of the oracle rollout within 1e-3; checkpointed and full-tape
gradients are identical; d v / d f_ext = h / m for a free particle
(reference test_solver.py:319-331).
"""

import numpy as np
import pytest
import torch

import oracle
from helpers import cube
from newton_solver_fem_amd import ModelBuilder, build_tet_model, meshes
from newton_solver_fem_amd.differentiation import rollout

pytestmark = pytest.mark.gpu

DEV = "cuda:0"
DT = 0.002
GROUND = dict(ground_plane=(0.0, 0.0, 1.0, -0.302), ground_ke=40.0, ground_kd=0.2, ground_kf=1.0, ground_mu=0.5)


def rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30))


def np_(t):
    return t.detach().cpu().numpy().astype(np.float64)


def make(kind, **extra):
    from newton_solver_fem_amd.solver import SolverFEM

    model = cube(kind, material=(500.0, 1000.0, 0.2), gravity=(0.0, 0.0, -9.81), device=DEV)
    kw = dict(dt=DT, min_stretch=-1.0, linear_damping=0.3, enable_tet_activations=(kind == "tet"), **GROUND)
    kw.update(extra)
    s = SolverFEM(model, solver_type="cg", preconditioner_type="block_diag", self_contact_ke=0.0, differentiable=True,
                  linear_solver_maxiter=128, **kw)
    names = dict(dt="dt", min_stretch="min_stretch", linear_damping="linear_damping",
                 enable_tet_activations="enable_tet_activations", ground_plane="ground_plane", ground_ke="ground_ke",
                 ground_kd="ground_kd", ground_kf="ground_kf", ground_mu="ground_mu", max_stretch="max_stretch")
    o = oracle.OracleSolver(model.host, bits=64, gravity=model.gravity,
                            **{names[k]: v for k, v in kw.items() if k in names})
    return model, s, o


def initial(model, seed):
    rng = np.random.default_rng(seed)
    q0 = model.host["particle_q"].astype(np.float64).copy()
    q0[:, 0] *= 1.03
    v0 = np.tile((0.02, -0.01, 0.03), (len(q0), 1)) + rng.normal(0, 0.01, q0.shape)
    return rng, q0, v0


@pytest.mark.parametrize("kind", ["tet", "hex"])
def test_step_and_its_gradients_match_oracle(kind):
    print(f"case diff step {kind}", flush=True)
    model, s, o = make(kind)
    rng, q0, v0 = initial(model, 3)
    ext = rng.normal(0, 0.02, q0.shape)
    act = rng.uniform(0, 0.3, model.tet_count) if kind == "tet" else None
    qo, vo, dv = o.diff_step(q0, v0, ext, act)
    gq_out, gv_out = rng.normal(size=q0.shape), rng.normal(size=q0.shape)
    ref = o.diff_step_backward(q0, v0, dv, gq_out, gv_out, act)

    tq = torch.tensor(q0, dtype=torch.float32, device=DEV, requires_grad=True)
    tv = torch.tensor(v0, dtype=torch.float32, device=DEV, requires_grad=True)
    te = torch.tensor(ext, dtype=torch.float32, device=DEV, requires_grad=True)
    ctrl = model.control()
    if act is not None:
        ctrl.tet_activations = torch.tensor(act, dtype=torch.float32, device=DEV, requires_grad=True)
        s.tet_materials.requires_grad_(True)
    st_in = model.state()
    st_in.particle_q, st_in.particle_qd = tq, tv
    st_out = model.state()
    s.step(st_in, st_out, ctrl, None, DT, external_forces=te)
    assert rel(np_(st_out.particle_q), qo) < 1e-6
    assert rel(np_(st_out.particle_qd), vo) < 1e-4
    loss = (st_out.particle_q * torch.tensor(gq_out, dtype=torch.float32, device=DEV)).sum() + (
        st_out.particle_qd * torch.tensor(gv_out, dtype=torch.float32, device=DEV)).sum()
    loss.backward()
    res = s.check_differentiation()
    assert res["forward_relative_residual"] < 2e-5 and res["backward_relative_residual"] < 2e-5
    assert rel(np_(tq.grad), ref["q"]) < 1e-3
    assert rel(np_(tv.grad), ref["qd"]) < 1e-3
    # The reverse products alone (without the identity pass-through of the update).
    assert rel(np_(tq.grad) - gq_out, ref["q"] - gq_out) < 2e-3
    u = gv_out + DT * gq_out
    assert rel(np_(tv.grad) - u, ref["qd"] - u) < 2e-3
    assert rel(np_(te.grad), ref["ext"]) < 1e-3
    if act is not None:
        assert rel(np_(ctrl.tet_activations.grad), ref["act"]) < 1e-3
        for col in range(3):
            assert rel(np_(s.tet_materials.grad)[:, col], ref["mat"][:, col]) < 2e-3, col


def _loss_np(q, v, w):
    return float(np.sum(w * q) + 0.03 * np.sum(v * v))


@pytest.mark.parametrize("checkpoint_every", [0, 4])
def test_rollout_gradients_match_oracle_finite_differences(checkpoint_every):
    """12 substeps, per-step activations and forces; FD of the FP64 oracle rollout (tol 1e-3)."""
    print(f"case diff rollout k={checkpoint_every}", flush=True)
    model, s, o = make("tet", defer_differentiation_checks=True)
    rng, q0, v0 = initial(model, 5)
    steps = 12
    exts = rng.normal(0, 0.02, (steps, *q0.shape))
    acts = rng.uniform(0, 0.3, (steps, model.tet_count))
    n = len(q0)
    w = np.stack([np.arange(1, n + 1) / n, np.full(n, 0.13), np.full(n, -0.41)], axis=1)

    def oracle_loss(q, v, e, a):
        for k in range(steps):
            q, v, _ = o.diff_step(q, v, e[k], a[k])
        return _loss_np(q, v, w)

    tq = torch.tensor(q0, dtype=torch.float32, device=DEV, requires_grad=True)
    tv = torch.tensor(v0, dtype=torch.float32, device=DEV, requires_grad=True)
    te = torch.tensor(exts, dtype=torch.float32, device=DEV, requires_grad=True)
    ta = torch.tensor(acts, dtype=torch.float32, device=DEV, requires_grad=True)
    qT, vT = rollout(s, tq, tv, steps, activations=ta, external_forces=te, checkpoint_every=checkpoint_every)
    tw = torch.tensor(w, dtype=torch.float32, device=DEV)
    loss = (tw * qT).sum() + 0.03 * (vT * vT).sum()
    loss.backward()
    s.check_differentiation()
    assert abs(loss.item() - oracle_loss(q0, v0, exts, acts)) < 1e-5 * abs(loss.item()) + 1e-7
    # fp64 FD with a small step: ground contact and friction have kinks that a
    # 1e-4 m perturbation can cross within 12 substeps.
    eps = 1e-6
    for name, g, base in (("q0", tq.grad, q0), ("v0", tv.grad, v0), ("ext", te.grad, exts), ("act", ta.grad, acts)):
        d = rng.normal(size=base.shape)
        args = {"q0": [q0, v0, exts, acts], "v0": [q0, v0, exts, acts], "ext": [q0, v0, exts, acts],
                "act": [q0, v0, exts, acts]}[name]
        idx = ("q0", "v0", "ext", "act").index(name)
        plus, minus = list(args), list(args)
        plus[idx] = base + eps * d
        minus[idx] = base - eps * d
        fd = (oracle_loss(*plus) - oracle_loss(*minus)) / (2 * eps)
        exact = float(np.sum(np_(g) * d))
        assert abs(exact - fd) <= 1e-3 * abs(fd) + 1e-6, (name, exact, fd)


def test_checkpointed_equals_full_tape():
    """Reference test_bear_training.py:76-138: checkpointed recomputation == full tape."""
    print("case checkpoint == full tape", flush=True)
    model, s, _ = make("tet", defer_differentiation_checks=True)
    rng, q0, v0 = initial(model, 9)
    acts = torch.tensor(rng.uniform(0, 0.3, model.tet_count), dtype=torch.float32, device=DEV)
    grads = []
    for k in (0, 3, 5):
        a = acts.clone().requires_grad_(True)
        tq = torch.tensor(q0, dtype=torch.float32, device=DEV, requires_grad=True)
        tv = torch.tensor(v0, dtype=torch.float32, device=DEV)
        qT, vT = rollout(s, tq, tv, 15, activations=a, checkpoint_every=k)
        (qT[:, 2].sum() + (vT * vT).sum()).backward()
        grads.append((np_(a.grad), np_(tq.grad)))
    s.check_differentiation()
    for g in grads[1:]:
        np.testing.assert_array_equal(g[0], grads[0][0])
        np.testing.assert_array_equal(g[1], grads[0][1])


def test_velocity_derivative_of_external_force_is_h_over_m():
    """Reference test_solver.py:319-331: one free particle of mass 2, d v_x / d f_x = h / m."""
    print("case free particle", flush=True)
    from newton_solver_fem_amd.solver import SolverFEM

    b = ModelBuilder(gravity=(0.0, 0.0, 0.0))
    b.add_particle((0, 0, 1), (0, 0, 0), mass=2.0)
    model = b.finalize(DEV)
    s = SolverFEM(model, dt=DT, solver_type="cg", preconditioner_type="block_diag", self_contact_ke=0.0,
                  differentiable=True, linear_solver_maxiter=128)
    st_in, st_out = model.state(), model.state()
    ext = torch.zeros(1, 3, device=DEV, requires_grad=True)
    s.step(st_in, st_out, None, None, DT, external_forces=ext)
    st_out.particle_qd[0, 0].backward()
    np.testing.assert_allclose(np_(ext.grad), [[DT / 2.0, 0.0, 0.0]], atol=1e-8)


def test_bar_fixed_base_mu_activation_gradient_matches_oracle():
    """The differentiable-bar workload at reduced size: base face fixed (projected solves, an
    extension), activations scaling mu (mu_e = mu0_e (1 + a_e)), tip-target loss after 20
    substeps, checkpoint every 5; d loss / d a vs FP64 central differences of the oracle (1e-3)."""
    print("case fixed-base bar", flush=True)
    from newton_solver_fem_amd.solver import SolverFEM

    sc = meshes.tet_bar_scene(nx=12, ny=3, nz=3)
    model = build_tet_model(sc, device=DEV)
    dt, steps = 1 / 600, 20
    s = SolverFEM(model, dt=dt, solver_type="cg", preconditioner_type="block_diag", self_contact_ke=0.0,
                  differentiable=True, defer_differentiation_checks=True, linear_solver_maxiter=256,
                  fixed_particle_indices=sc.fixed, allow_fixed_in_differentiation=True)
    mat0 = torch.tensor(sc.materials, dtype=torch.float32, device=DEV)
    a = torch.tensor(sc.activations, dtype=torch.float32, device=DEV, requires_grad=True)
    scale = torch.stack([1.0 + a, torch.ones_like(a), torch.ones_like(a)], dim=1)
    s.set_tet_materials(mat0 * scale)
    rest = model.host["particle_q"].astype(np.float64)
    rest_t = model.particle_q.clone()
    qT, _ = rollout(s, rest_t.clone(), torch.zeros_like(rest_t), steps, checkpoint_every=5)
    tip = rest[:, 0] > rest[:, 0].max() - 1e-9
    tgt = rest[tip].mean(0) + np.array([0.004, -0.002, 0.003])
    loss = ((qT[torch.tensor(tip, device=DEV)].mean(0) - torch.tensor(tgt, dtype=torch.float32, device=DEV)) ** 2).sum()
    loss.backward()
    s.check_differentiation()
    assert np.all(np_(qT)[sc.fixed] == rest[sc.fixed].astype(np.float32))

    o = oracle.OracleSolver(model.host, bits=64, dt=dt, gravity=model.gravity, fixed=sc.fixed)

    def oracle_loss(acts):
        m = sc.materials.astype(np.float64).copy()
        m[:, 0] *= 1.0 + acts
        o.set_tet_materials(m)
        o.refresh_tangent(rest)
        q, v = rest.copy(), np.zeros_like(rest)
        for _ in range(steps):
            q, v, _ = o.diff_step(q, v)
        return float(np.sum((q[tip].mean(0) - tgt) ** 2))

    a0 = sc.activations.astype(np.float64)
    assert abs(loss.item() - oracle_loss(a0)) <= 1e-4 * loss.item() + 1e-12
    rng = np.random.default_rng(4)
    d = rng.normal(size=model.tet_count)
    eps = 1e-6
    fd = (oracle_loss(a0 + eps * d) - oracle_loss(a0 - eps * d)) / (2 * eps)
    exact = float(np.sum(np_(a.grad) * d))
    assert abs(exact - fd) <= 1e-3 * abs(fd) + 1e-12, (exact, fd)


def test_step_gradients_with_springs_match_oracle():
    """Tet4 cube + damped edge springs + damped triangle membranes: one
    differentiable step and its input gradients vs the oracle's reverse products."""
    print("case diff springs", flush=True)
    from newton_solver_fem_amd.solver import SolverFEM

    base = cube("tet", material=(500.0, 1000.0, 0.2), gravity=(0.0, 0.0, -9.81))
    b = ModelBuilder(gravity=(0.0, 0.0, -9.81))
    b.add_particles(base.host["particle_q"], np.zeros((base.particle_count, 3)), base.host["particle_mass"],
                    radius=0.004)
    b.add_tetrahedra(base.host["tet_indices"], base.host["tet_materials"])
    tets = base.host["tet_indices"].tolist()
    edges = sorted({(min(a, c), max(a, c)) for t in tets for a, c in ((t[0], t[1]), (t[1], t[2]), (t[2], t[3]))})
    b.add_springs(edges, 30.0, 0.4)
    b.add_triangles([t[:3] for t in tets[:6]], tri_ke=30.0, tri_ka=50.0, tri_kd=0.2)
    model = b.finalize(DEV)
    kw = dict(dt=DT, min_stretch=-1.0, linear_damping=0.3, **GROUND)
    s = SolverFEM(model, solver_type="cg", preconditioner_type="block_diag", self_contact_ke=0.0, differentiable=True,
                  linear_solver_maxiter=128, **kw)
    o = oracle.OracleSolver(model.host, bits=64, gravity=model.gravity, **kw)
    rng, q0, v0 = initial(model, 5)
    ext = rng.normal(0, 0.02, q0.shape)
    qo, vo, dv = o.diff_step(q0, v0, ext)
    gq_out, gv_out = rng.normal(size=q0.shape), rng.normal(size=q0.shape)
    ref = o.diff_step_backward(q0, v0, dv, gq_out, gv_out)
    tq = torch.tensor(q0, dtype=torch.float32, device=DEV, requires_grad=True)
    tv = torch.tensor(v0, dtype=torch.float32, device=DEV, requires_grad=True)
    te = torch.tensor(ext, dtype=torch.float32, device=DEV, requires_grad=True)
    st_in, st_out = model.state(), model.state()
    st_in.particle_q, st_in.particle_qd = tq, tv
    s.step(st_in, st_out, None, None, DT, external_forces=te)
    assert rel(np_(st_out.particle_q), qo) < 1e-6
    assert rel(np_(st_out.particle_qd), vo) < 1e-4
    loss = (st_out.particle_q * torch.tensor(gq_out, dtype=torch.float32, device=DEV)).sum() + (
        st_out.particle_qd * torch.tensor(gv_out, dtype=torch.float32, device=DEV)).sum()
    loss.backward()
    u = gv_out + DT * gq_out
    assert rel(np_(tq.grad) - gq_out, ref["q"] - gq_out) < 2e-3
    assert rel(np_(tv.grad) - u, ref["qd"] - u) < 2e-3
    assert rel(np_(te.grad), ref["ext"]) < 1e-3


def _two_step_loss(model, s, q0, v0, act):
    """Two recorded steps from (q0, v0) with activations `act`; returns the
    loss and the leaf tensors."""
    q = torch.tensor(q0, dtype=torch.float32, device=DEV, requires_grad=True)
    v = torch.tensor(v0, dtype=torch.float32, device=DEV, requires_grad=True)
    a = torch.tensor(act, dtype=torch.float32, device=DEV, requires_grad=True)
    ctrl = model.control()
    ctrl.tet_activations = a
    s0, s1, s2 = model.state(), model.state(), model.state()
    s0.particle_q, s0.particle_qd = q, v
    s.step(s0, s1, ctrl, None, DT)
    s.step(s1, s2, ctrl, None, DT)
    loss = (s2.particle_q * s2.particle_q).sum() + 0.1 * (s2.particle_qd * s2.particle_qd).sum()
    return loss, (q, v, a)


def test_recorded_backward_survives_unrelated_forward_and_rejects_revision_change():
    """Reference tests/test_activation_gradients.py:147-166 and :227-233: a
    recorded rollout owns every primal its backward needs (unrelated forward
    steps in between leave its gradients unchanged), and reassembling the
    matrix after recording invalidates the backward."""
    print("case diff snapshots", flush=True)
    model, s, _ = make("tet")
    _, q0, v0 = initial(model, 5)
    act = np.full(model.tet_count, 0.03, np.float32)
    s.refresh_tangent(torch.tensor(q0, dtype=torch.float32, device=DEV))
    loss, leaves = _two_step_loss(model, s, q0, v0, act)
    with torch.no_grad():  # unrelated steps overwrite every persistent solver buffer
        a, b = model.state(), model.state()
        a.particle_q.copy_(torch.tensor(q0 + (0.01, 0.02, 0.0), dtype=torch.float32))
        a.particle_qd.copy_(torch.tensor(-v0, dtype=torch.float32))
        ctrl = model.control()
        ctrl.tet_activations = torch.full_like(ctrl.tet_activations, -0.03)
        for _ in range(3):
            s.step(a, b, ctrl, None, DT)
            a, b = b, a
    loss.backward()
    first = [t.grad.clone() for t in leaves]
    loss2, leaves2 = _two_step_loss(model, s, q0, v0, act)
    loss2.backward()
    for g1, t2 in zip(first, leaves2):
        np.testing.assert_allclose(np_(g1), np_(t2.grad), rtol=2e-6, atol=5e-6)
    loss3, _ = _two_step_loss(model, s, q0, v0, act)
    s.refresh_tangent(torch.tensor(q0, dtype=torch.float32, device=DEV))
    with pytest.raises(RuntimeError, match="changed"):
        loss3.backward()


def test_taped_step_rejects_aliased_states():
    """Reference tests/test_activation_gradients.py:211-225: a recorded step
    needs distinct input and output states."""
    print("case diff alias", flush=True)
    model, s, _ = make("tet")
    _, q0, v0 = initial(model, 6)
    cur = model.state()
    cur.particle_q = torch.tensor(q0, dtype=torch.float32, device=DEV, requires_grad=True)
    cur.particle_qd = torch.tensor(v0, dtype=torch.float32, device=DEV, requires_grad=True)
    with pytest.raises(ValueError, match="unique"):
        s.step(cur, cur, None, None, DT)


def test_forward_is_independent_of_interleaved_large_adjoint_solves():
    """Reference tests/test_solver.py:185-201: adjoint solves with huge seeds
    between two forward rollouts leave the forward result unchanged."""
    print("case diff interleaved adjoint", flush=True)
    model, s, _ = make("tet")
    _, q0, v0 = initial(model, 7)
    act = np.full(model.tet_count, 0.02, np.float32)
    s.refresh_tangent(torch.tensor(q0, dtype=torch.float32, device=DEV))

    def forward():
        with torch.no_grad():
            loss, _ = _two_step_loss(model, s, q0, v0, act)
        return float(loss)

    baseline = forward()
    for seed in (1.0e4, 1.0e8):
        loss, _ = _two_step_loss(model, s, q0, v0, act)
        (loss * seed).backward()
        assert forward() == pytest.approx(baseline, rel=1e-7, abs=0)


def test_step_gradients_with_triangle_drag_and_lift_match_finite_differences():
    """Triangle membranes with drag and lift (eval_triangles kernels.py:876-970): the
    gradient of a one-step loss w.r.t. the input positions and velocities vs FP64
    central differences of the oracle's differentiable step (fixed matrix), along
    random directions; the oracle's forward includes the aerodynamic terms."""
    print("case diff drag lift", flush=True)
    from newton_solver_fem_amd.solver import SolverFEM

    b = ModelBuilder(gravity=(0.0, 0.0, -9.81))
    n = 5
    g = np.stack(np.meshgrid(np.arange(n), np.arange(n), indexing="ij"), -1).reshape(-1, 2) * 0.2
    pts = np.column_stack([g, np.zeros(len(g)) + 0.3])
    b.add_particles(pts, np.zeros_like(pts), 0.002, radius=0.0)
    tris = []
    for i in range(n - 1):
        for j in range(n - 1):
            a0, a1, a2, a3 = i * n + j, (i + 1) * n + j, (i + 1) * n + j + 1, i * n + j + 1
            tris += [(a0, a1, a2), (a0, a2, a3)]
    # soft, large triangles: drag and lift carry 12-80% of the force terms' gradient here
    b.add_triangles(tris, tri_ke=2.0, tri_ka=2.0, tri_kd=0.05, tri_drag=3.0, tri_lift=5.0)
    model = b.finalize(DEV)
    kw = dict(dt=DT, linear_damping=0.2)
    s = SolverFEM(model, solver_type="cg", preconditioner_type="block_diag", self_contact_ke=0.0, differentiable=True,
                  linear_solver_maxiter=128, **kw)
    o = oracle.OracleSolver(model.host, bits=64, gravity=model.gravity, **kw)
    rng = np.random.default_rng(17)
    q0 = model.host["particle_q"].astype(np.float64) + rng.normal(0, 0.002, pts.shape)
    v0 = rng.normal(0, 0.5, pts.shape) + (0.3, -0.2, 0.6)  # wind across and through the sheet
    gq_out, gv_out = rng.normal(size=pts.shape), rng.normal(size=pts.shape)
    tq = torch.tensor(q0, dtype=torch.float32, device=DEV, requires_grad=True)
    tv = torch.tensor(v0, dtype=torch.float32, device=DEV, requires_grad=True)
    st_in, st_out = model.state(), model.state()
    st_in.particle_q, st_in.particle_qd = tq, tv
    s.step(st_in, st_out, None, None, DT)
    qo, vo, _ = o.diff_step(q0, v0)
    assert rel(np_(st_out.particle_qd), vo) < 1e-4
    loss = (st_out.particle_q * torch.tensor(gq_out, dtype=torch.float32, device=DEV)).sum() + (
        st_out.particle_qd * torch.tensor(gv_out, dtype=torch.float32, device=DEV)).sum()
    loss.backward()
    gq, gv = np_(tq.grad), np_(tv.grad)

    def L(q, v):
        a, c, _ = o.diff_step(q, v)
        return float(np.sum(a * gq_out) + np.sum(c * gv_out))

    eps = 1e-6
    for k in range(3):
        dq, dvd = rng.normal(size=pts.shape), rng.normal(size=pts.shape)
        fd_q = (L(q0 + eps * dq, v0) - L(q0 - eps * dq, v0)) / (2 * eps)
        fd_v = (L(q0, v0 + eps * dvd) - L(q0, v0 - eps * dvd)) / (2 * eps)
        # subtract the identity pass-through (q_out = q_in + h v_out) so the force terms are compared
        base_q, base_v = float(np.sum(gq_out * dq)), float(np.sum((gv_out + DT * gq_out) * dvd))
        assert abs(float(np.sum(gq * dq)) - fd_q) <= 2e-3 * max(abs(fd_q - base_q), 1e-12) + 1e-6 * abs(fd_q), k
        assert abs(float(np.sum(gv * dvd)) - fd_v) <= 2e-3 * max(abs(fd_v - base_v), 1e-12) + 1e-6 * abs(fd_v), k


def test_differentiation_refuses_particle_contact_stiffness():
    """Reference prepare (differentiation.py:238-242): a model whose
    particle_ke (SolverFEM's self_contact_ke) is positive cannot be
    differentiated, grid or not."""
    print("case diff self_contact_ke", flush=True)
    from newton_solver_fem_amd.solver import SolverFEM

    model = cube("tet", density=100.0, gravity=(0.0, 0.0, -9.81), device=DEV)
    s = SolverFEM(model, solver_type="cg", preconditioner_type="block_diag", differentiable=True, dt=DT)
    a, b = model.state(), model.state()
    with pytest.raises(NotImplementedError, match="analytic ground contact"):
        s.step(a, b, None, None, DT)
    with pytest.raises(NotImplementedError, match="analytic ground contact"):
        s.prepare_differentiation()
