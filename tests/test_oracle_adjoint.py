"""Pins the FP64 oracle's reverse-mode products (the checker of the HIP adjoint path).

Every VJP of the oracle is compared with central finite differences of the
oracle's own forward functions, which the other oracle tests pin to the
reference (energy gradients, closed forms, dense solves):

* per force term (Tet4 incl. activations and the stretch clamp engaged,
  Hex8, ground contact with friction, linear damping) w.r.t. positions,
  velocities, activations and Tet4 materials;
* the whole differentiable step (differentiation.py: fixed matrix, checked
  solve, predictor, update) over a 3-step rollout w.r.t. initial state,
  external forces, activations and Tet4 materials (mu, lambda, with the
  matrix re-assembled from the perturbed materials -- the terms the
  reference leaves out).
Tolerances: rtol 1e-6 for element terms, 2e-6 for rollouts (fp64 central
differences with eps 1e-6 and solves to 1e-13).
"""

import numpy as np
import pytest

import oracle
from helpers import cube

T_TET, T_HEX, T_GROUND, T_DAMP = oracle.TERM_TET, oracle.TERM_HEX, oracle.TERM_GROUND, oracle.TERM_DAMPING


def _state(model, rng, squeeze=0.0):
    q = model.host["particle_q"].astype(np.float64).copy()
    c = q.mean(axis=0)
    q = c + (q - c) * np.array([1.0 - squeeze, 1.0 + 0.5 * squeeze, 1.0])
    q += rng.normal(0.0, 0.004, q.shape)
    v = rng.normal(0.0, 0.05, q.shape)
    return q, v


def _dd(f, x, d, eps=1e-6):
    return (f(x + eps * d) - f(x - eps * d)) / (2 * eps)


def _check(exact, numeric, rtol, what):
    scale = max(abs(numeric), abs(exact), 1e-12)
    assert abs(exact - numeric) <= rtol * scale + 1e-12, f"{what}: exact {exact!r} numeric {numeric!r}"


@pytest.mark.parametrize("case", ["tet", "tet_clamped", "hex", "hex_clamped", "ground", "damping"])
def test_term_vjp_matches_finite_differences(case):
    rng = np.random.default_rng(hash(case) % 2**32)
    kind = "hex" if case.startswith("hex") else "tet"
    model = cube(kind, material=(500.0, 1000.0, 0.3), gravity=(0.0, 0.0, -9.81))
    kw = dict(enable_tet_activations=True)
    if case.endswith("clamped"):
        kw.update(min_stretch=0.93, max_stretch=1.07)
    if case == "ground":
        kw.update(ground_plane=(0.0, 0.0, 1.0, -0.305), ground_ke=50.0, ground_kd=0.3, ground_kf=2.0, ground_mu=0.4)
    if case == "damping":
        kw.update(linear_damping=0.7)
    o = oracle.OracleSolver(model.host, bits=64, gravity=model.gravity, **kw)
    mask = {"tet": T_TET, "tet_clamped": T_TET, "hex": T_HEX, "hex_clamped": T_HEX, "ground": T_GROUND,
            "damping": T_DAMP}[case]
    q, v = _state(model, rng, squeeze=0.15 if case.endswith("clamped") else 0.02)
    if case == "ground":  # half of the nodes below the plane, tangential slip in both friction regimes
        v[:, :2] *= 10.0
    act = rng.uniform(0.0, 0.3, o.T) if o.T else None
    z = rng.normal(size=q.shape)

    def L(qq, vv, aa=act):
        return float(np.sum(z * o.eval_terms(mask, qq, vv, aa)))

    xb, vb, ab, mb = o.vjp_terms(mask, q, v, z, act)
    # Smaller steps where the clamp is engaged: a singular value near a bound
    # puts a kink within reach of a 1e-6 step.
    eps = 1e-7 if case.endswith("clamped") else 1e-6
    for k in range(2):
        d = rng.normal(size=q.shape)
        _check(float(np.sum(xb * d)), _dd(lambda x: L(x, v), q, d, eps), 1e-6, f"{case} dq")
        _check(float(np.sum(vb * d)), _dd(lambda x: L(q, x), v, d, eps), 1e-6, f"{case} dv")
    if mask == T_TET:
        d = rng.normal(size=o.T)
        _check(float(np.sum(ab * d)), _dd(lambda a: L(q, v, a), act, d), 1e-6, "activation")
        mat0 = model.host["tet_materials"].astype(np.float64)
        for col in range(3):
            d = np.zeros_like(mat0)
            d[:, col] = rng.normal(size=o.T) * mat0[:, col].mean()

            def Lm(m):
                o.set_tet_materials(m)
                return L(q, v)

            num = _dd(Lm, mat0, d, eps=1e-5)
            o.set_tet_materials(mat0)
            _check(float(np.sum(mb * d)), num, 1e-6, f"material {col}")
    if case.endswith("clamped"):  # the clamp must actually be engaged somewhere
        assert np.any(np.abs(o.eval_terms(mask, q, v) - oracle.OracleSolver(
            model.host, bits=64, gravity=model.gravity, enable_tet_activations=True).eval_terms(mask, q, v)) > 1e-9)


def _rollout(o, q, v, exts, acts, weights):
    states = [(q, v)]
    dvs = []
    for k in range(len(exts)):
        qn, vn, dv = o.diff_step(states[-1][0], states[-1][1], exts[k], None if acts is None else acts[k])
        states.append((qn, vn))
        dvs.append(dv)
    qT, vT = states[-1]
    loss = float(np.sum(weights * qT) + 0.03 * np.sum(vT * vT))
    return loss, states, dvs


def _rollout_grad(o, states, dvs, exts, acts, weights):
    qb, vb = weights.copy(), 0.06 * states[-1][1]
    g_ext, g_act, g_mat = [None] * len(exts), [None] * len(exts), 0.0
    for k in reversed(range(len(exts))):
        q, v = states[k]
        out = o.diff_step_backward(q, v, dvs[k], qb, vb, None if acts is None else acts[k])
        qb, vb = out["q"], out["qd"]
        g_ext[k], g_act[k] = out["ext"], out["act"]
        g_mat = g_mat + out["mat"]
    return qb, vb, g_ext, g_act, g_mat


@pytest.mark.parametrize("kind", ["tet", "hex"])
def test_differentiable_rollout_gradients_match_finite_differences(kind):
    """Reference test_solver.py:264-305 (multistep state/force gradients), in FP64 on the oracle."""
    rng = np.random.default_rng(7 if kind == "tet" else 8)
    model = cube(kind, material=(500.0, 1000.0, 0.2), gravity=(0.0, 0.0, -9.81))
    o = oracle.OracleSolver(model.host, bits=64, dt=0.002, gravity=model.gravity, min_stretch=-1.0,
                            enable_tet_activations=(kind == "tet"), linear_damping=0.3,
                            ground_plane=(0.0, 0.0, 1.0, -0.302), ground_ke=40.0, ground_kd=0.2, ground_kf=1.0,
                            ground_mu=0.5)
    q0 = model.host["particle_q"].astype(np.float64).copy()
    q0[:, 0] *= 1.03
    v0 = np.tile((0.02, -0.01, 0.03), (len(q0), 1)) + rng.normal(0, 0.01, q0.shape)
    steps = 3
    exts = [rng.normal(0, 0.02, q0.shape) for _ in range(steps)]
    acts = [rng.uniform(0, 0.3, o.T) for _ in range(steps)] if kind == "tet" else None
    n = len(q0)
    weights = np.stack([np.arange(1, n + 1) / n, np.full(n, 0.13), np.full(n, -0.41)], axis=1)
    base, states, dvs = _rollout(o, q0, v0, exts, acts, weights)
    gq, gv, g_ext, g_act, g_mat = _rollout_grad(o, states, dvs, exts, acts, weights)

    d = rng.normal(size=q0.shape)
    _check(float(np.sum(gq * d)), _dd(lambda x: _rollout(o, x, v0, exts, acts, weights)[0], q0, d), 2e-6, "q0")
    _check(float(np.sum(gv * d)), _dd(lambda x: _rollout(o, q0, x, exts, acts, weights)[0], v0, d), 2e-6, "v0")
    k = 1

    def with_ext(x):
        e = list(exts)
        e[k] = x
        return _rollout(o, q0, v0, e, acts, weights)[0]

    _check(float(np.sum(g_ext[k] * d)), _dd(with_ext, exts[k], d), 2e-6, "ext")
    if kind != "tet":
        return
    da = rng.normal(size=o.T)

    def with_act(x):
        a = list(acts)
        a[k] = x
        return _rollout(o, q0, v0, exts, a, weights)[0]

    _check(float(np.sum(g_act[k] * da)), _dd(with_act, acts[k], da), 2e-6, "activation")
    mat0 = model.host["tet_materials"].astype(np.float64)
    rest = model.host["particle_q"].astype(np.float64)
    for col in (0, 1):
        dm = np.zeros_like(mat0)
        dm[:, col] = rng.normal(size=o.T) * mat0[:, col].mean()

        def with_mat(m):
            o.set_tet_materials(m)
            o.refresh_tangent(rest)
            return _rollout(o, q0, v0, exts, acts, weights)[0]

        num = _dd(with_mat, mat0, dm, eps=1e-5)
        o.set_tet_materials(mat0)
        o.refresh_tangent(rest)
        _check(float(np.sum(g_mat * dm)), num, 2e-6, f"material {col} (with the matrix terms)")


def test_velocity_derivative_of_external_force_is_h_over_m():
    """Reference test_solver.py:319-331: one free particle, d v_x / d f_ext,x = h / m."""
    from newton_solver_fem_amd import ModelBuilder

    b = ModelBuilder(gravity=(0.0, 0.0, 0.0))
    b.add_particle((0, 0, 1), (0, 0, 0), mass=2.0)
    model = b.finalize("cpu")
    o = oracle.OracleSolver(model.host, bits=64, dt=0.002, gravity=(0.0, 0.0, 0.0))
    q, v = model.host["particle_q"].astype(np.float64), np.zeros((1, 3))
    _, _, dv = o.diff_step(q, v, np.zeros((1, 3)))
    gv = np.array([[1.0, 0.0, 0.0]])
    out = o.diff_step_backward(q, v, dv, np.zeros((1, 3)), gv)
    np.testing.assert_allclose(out["ext"], [[0.002 / 2.0, 0.0, 0.0]], atol=1e-15)


def test_differentiable_rollout_with_fixed_particles():
    """Extension (the reference rejects fixed particles when differentiable):
    projected forward/adjoint solves; gradients of a 3-step rollout vs FD."""
    from newton_solver_fem_amd import build_tet_model, meshes

    rng = np.random.default_rng(12)
    sc = meshes.tet_bar_scene(nx=6, ny=2, nz=2)
    sc.materials = sc.materials.copy()
    sc.materials[:, 2] = 0.05
    model = build_tet_model(sc, device="cpu")
    o = oracle.OracleSolver(model.host, bits=64, dt=1 / 600, gravity=model.gravity, fixed=sc.fixed,
                            enable_tet_activations=True, linear_damping=0.2)
    rest = model.host["particle_q"].astype(np.float64)
    q0 = rest + rng.normal(0, 1e-4, rest.shape)
    v0 = rng.normal(0, 0.02, rest.shape)
    steps = 3
    exts = [rng.normal(0, 0.01, rest.shape) for _ in range(steps)]
    acts = [rng.uniform(0, 0.3, o.T) for _ in range(steps)]
    n = len(rest)
    weights = rng.normal(size=(n, 3))
    base, states, dvs = _rollout(o, q0, v0, exts, acts, weights)
    for q, v in states[1:]:
        np.testing.assert_array_equal(q[sc.fixed], rest[sc.fixed])
        np.testing.assert_array_equal(v[sc.fixed], 0.0)
    gq, gv, g_ext, g_act, g_mat = _rollout_grad(o, states, dvs, exts, acts, weights)
    assert np.all(gq[sc.fixed] == 0.0) and np.all(gv[sc.fixed] == 0.0)
    d = rng.normal(size=rest.shape)
    _check(float(np.sum(gq * d)), _dd(lambda x: _rollout(o, x, v0, exts, acts, weights)[0], q0, d), 2e-6, "q0")
    _check(float(np.sum(gv * d)), _dd(lambda x: _rollout(o, q0, x, exts, acts, weights)[0], v0, d), 2e-6, "v0")
    da = rng.normal(size=o.T)

    def with_act(x):
        a = list(acts)
        a[1] = x
        return _rollout(o, q0, v0, exts, a, weights)[0]

    _check(float(np.sum(g_act[1] * da)), _dd(with_act, acts[1], da), 2e-6, "activation")
    mat0 = model.host["tet_materials"].astype(np.float64)
    dm = np.zeros_like(mat0)
    dm[:, 0] = rng.normal(size=o.T) * mat0[:, 0].mean()

    def with_mat(m):
        o.set_tet_materials(m)
        o.refresh_tangent(rest)
        return _rollout(o, q0, v0, exts, acts, weights)[0]

    num = _dd(with_mat, mat0, dm, eps=1e-5)
    o.set_tet_materials(mat0)
    o.refresh_tangent(rest)
    _check(float(np.sum(g_mat * dm)), num, 2e-6, "mu with fixed particles")


def test_spring_vjp_matches_finite_differences():
    """Reverse product of eval_springs (kernels.py:176-219) vs central differences."""
    from newton_solver_fem_amd import ModelBuilder

    rng = np.random.default_rng(31)
    b = ModelBuilder(gravity=(0.0, 0.0, 0.0))
    for p in rng.uniform(-0.1, 0.1, (6, 3)):
        b.add_particle(p, (0, 0, 0), mass=0.01)
    for i, j, k, c in ((0, 1, 40.0, 0.3), (1, 2, 25.0, 0.0), (2, 3, 60.0, 0.7), (3, 4, 10.0, 0.2), (0, 5, 35.0, 0.1)):
        b.add_spring(i, j, k, c)
    model = b.finalize("cpu")
    o = oracle.OracleSolver(model.host, bits=64, gravity=model.gravity)
    q = model.host["particle_q"].astype(np.float64) + rng.normal(0, 0.01, (6, 3))
    v = rng.normal(0, 0.1, q.shape)
    z = rng.normal(size=q.shape)

    def L(qq, vv):
        return float(np.sum(z * o.eval_terms(oracle.TERM_SPRING, qq, vv)))

    xb, vb, _, _ = o.vjp_terms(oracle.TERM_SPRING, q, v, z)
    for _ in range(2):
        d = rng.normal(size=q.shape)
        _check(float(np.sum(xb * d)), _dd(lambda x: L(x, v), q, d), 1e-6, "spring dq")
        _check(float(np.sum(vb * d)), _dd(lambda x: L(q, x), v, d), 1e-6, "spring dv")


def test_differentiable_rollout_with_springs():
    """Tet4 cube plus edge springs (with damping, so the spring-damping predictor
    restore enters the step): 3-step rollout gradients vs FD."""
    rng = np.random.default_rng(33)
    model = cube("tet", material=(500.0, 1000.0, 0.2), gravity=(0.0, 0.0, -9.81))
    from newton_solver_fem_amd import ModelBuilder

    b = ModelBuilder(gravity=(0.0, 0.0, -9.81))
    b.add_particles(model.host["particle_q"], np.zeros((model.particle_count, 3)), model.host["particle_mass"],
                    radius=0.004)
    b.add_tetrahedra(model.host["tet_indices"], model.host["tet_materials"])
    edges = sorted({(min(a, c), max(a, c)) for t in model.host["tet_indices"].tolist()
                    for a, c in ((t[0], t[1]), (t[1], t[2]), (t[2], t[3]), (t[0], t[3]))})
    b.add_springs(edges, 30.0, 0.4)
    model = b.finalize("cpu")
    o = oracle.OracleSolver(model.host, bits=64, dt=0.002, gravity=model.gravity, min_stretch=-1.0,
                            linear_damping=0.3)
    q0 = model.host["particle_q"].astype(np.float64).copy()
    q0[:, 0] *= 1.03
    v0 = np.tile((0.02, -0.01, 0.03), (len(q0), 1)) + rng.normal(0, 0.01, q0.shape)
    exts = [rng.normal(0, 0.02, q0.shape) for _ in range(3)]
    n = len(q0)
    weights = np.stack([np.arange(1, n + 1) / n, np.full(n, 0.13), np.full(n, -0.41)], axis=1)
    base, states, dvs = _rollout(o, q0, v0, exts, None, weights)
    gq, gv, g_ext, _, _ = _rollout_grad(o, states, dvs, exts, None, weights)
    d = rng.normal(size=q0.shape)
    _check(float(np.sum(gq * d)), _dd(lambda x: _rollout(o, x, v0, exts, None, weights)[0], q0, d), 2e-6, "q0")
    _check(float(np.sum(gv * d)), _dd(lambda x: _rollout(o, q0, x, exts, None, weights)[0], v0, d), 2e-6, "v0")

    def with_ext(x):
        e = list(exts)
        e[1] = x
        return _rollout(o, q0, v0, e, None, weights)[0]

    _check(float(np.sum(g_ext[1] * d)), _dd(with_ext, exts[1], d), 2e-6, "ext")


def test_triangle_vjp_matches_finite_differences():
    """Reverse product of eval_triangles (membrane, area preservation, damping;
    no drag or lift) vs central differences."""
    from newton_solver_fem_amd import ModelBuilder

    rng = np.random.default_rng(35)
    b = ModelBuilder(gravity=(0.0, 0.0, 0.0))
    for p in ((0, 0, 0), (0.1, 0, 0), (0.0, 0.12, 0.01), (0.11, 0.1, 0.03), (0.05, -0.08, 0.02)):
        b.add_particle(p, (0, 0, 0), mass=0.01)
    b.add_triangle(0, 1, 2, tri_ke=30.0, tri_ka=50.0, tri_kd=0.4)
    b.add_triangle(1, 3, 2, tri_ke=20.0, tri_ka=70.0, tri_kd=0.2)
    b.add_triangle(0, 4, 1, tri_ke=25.0, tri_ka=40.0, tri_kd=0.3)
    model = b.finalize("cpu")
    o = oracle.OracleSolver(model.host, bits=64, gravity=model.gravity)
    q = model.host["particle_q"].astype(np.float64) + rng.normal(0, 0.01, (5, 3))
    v = rng.normal(0, 0.1, q.shape)
    z = rng.normal(size=q.shape)

    def L(qq, vv):
        return float(np.sum(z * o.eval_terms(oracle.TERM_TRI, qq, vv)))

    xb, vb, _, _ = o.vjp_terms(oracle.TERM_TRI, q, v, z)
    for _ in range(3):
        d = rng.normal(size=q.shape)
        _check(float(np.sum(xb * d)), _dd(lambda x: L(x, v), q, d), 1e-6, "triangle dq")
        _check(float(np.sum(vb * d)), _dd(lambda x: L(q, x), v, d), 1e-6, "triangle dv")
