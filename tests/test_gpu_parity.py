"""HIP library vs the CPU oracle on identical inputs (C-ABI path on cuda:0).

Tolerances: BSR offsets/columns and fixed masks are compared bit-exactly;
float32 GPU results are compared with the FP64 oracle relative to the scale
of each quantity (reduction order differs).  Positions after multi-substep
rollouts must agree within 1e-4 relative.
"""

import numpy as np
import pytest
import torch

import oracle
from helpers import brute_pattern, cube, tet_pair
from newton_solver_fem_amd import build_hex_model, build_tet_model, meshes

pytestmark = pytest.mark.gpu

DEV = "cuda:0"


def solver_for(model, **kw):
    from newton_solver_fem_amd.solver import SolverFEM

    base = dict(solver_type="cg", preconditioner_type="block_diag", self_contact_ke=0.0)
    base.update(kw)
    return SolverFEM(model, **base)


def oracle_for(model, bits=64, fixed=(), **kw):
    names = dict(dt="dt", min_stretch="min_stretch", max_stretch="max_stretch", ground_plane="ground_plane",
                 ground_ke="ground_ke", ground_kd="ground_kd", ground_kf="ground_kf", ground_mu="ground_mu",
                 linear_damping="linear_damping", enable_tet_activations="enable_tet_activations",
                 linear_solver_tolerance="tolerance", linear_solver_maxiter="maxiter")
    args = {names[k]: v for k, v in kw.items() if k in names}
    return oracle.OracleSolver(model.host, bits=bits, gravity=model.gravity, fixed=fixed, **args)


def rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30))


def np_(t):
    return t.detach().cpu().numpy().astype(np.float64)


# ---------------------------------------------------------------------------
@pytest.mark.parametrize("kind", ["tet", "hex"])
def test_pattern_bit_exact_small(kind):
    print(f"case pattern {kind}", flush=True)
    model = cube(kind, device=DEV)
    s = solver_for(model, dt=0.002, min_stretch=-1.0)
    ref = oracle_for(model, dt=0.002, min_stretch=-1.0)
    off, cols, _, _ = ref.matrix()
    np.testing.assert_array_equal(s.A_bsr.offsets.cpu().numpy(), off)
    np.testing.assert_array_equal(s.A_bsr.columns.cpu().numpy(), cols)
    if kind == "tet":
        assert s.A_bsr.nnz == 207  # tet_grid.usda: 27 nodes, 40 tets


@pytest.mark.parametrize("config", ["drop", "cantilever", "bar", "large"])
def test_pattern_bit_exact_configs(config):
    print(f"case pattern {config}", flush=True)
    if config == "cantilever":
        sc = meshes.hex_cantilever_scene()
        model = build_hex_model(sc, device=DEV)
        off, cols = brute_pattern(len(sc.points), hexes=sc.hexes)
        expect_nnz = 463_393
    else:
        sc = {"drop": meshes.tet_drop_scene, "bar": meshes.tet_bar_scene, "large": meshes.large_tet_scene}[config]()
        model = build_tet_model(sc, device=DEV)
        off, cols = brute_pattern(len(sc.points), tets=sc.tets)
        expect_nnz = {"drop": 15_191, "bar": 38_377, "large": 2_693_569}[config]
    s = solver_for(model, dt=1 / 600, fixed_particle_indices=sc.fixed)
    assert s.A_bsr.nnz == expect_nnz
    np.testing.assert_array_equal(s.A_bsr.offsets.cpu().numpy(), off)
    np.testing.assert_array_equal(s.A_bsr.columns.cpu().numpy(), cols)


@pytest.mark.parametrize("kind", ["tet", "hex"])
def test_assembly_matches_oracle(kind):
    print(f"case assembly {kind}", flush=True)
    model = cube(kind, device=DEV)
    rng = np.random.default_rng(3)
    q = model.host["particle_q"].astype(np.float64) + rng.normal(0, 0.004, (model.particle_count, 3))
    s = solver_for(model, dt=0.002)
    ref = oracle_for(model, dt=0.002)
    qt = torch.tensor(q, dtype=torch.float32, device=DEV)
    s.refresh_tangent(qt)
    qh = qt.cpu().numpy().astype(np.float64)
    ref.refresh_tangent(qh)
    _, _, vals, minv = ref.matrix()
    gv = np_(s.A_bsr.values)
    assert rel(gv, vals) < 2e-6
    assert rel(np_(s._pc_inv_diag), minv) < 2e-5


@pytest.mark.parametrize("kind", ["tet", "hex"])
def test_forces_match_oracle(kind):
    print(f"case forces {kind}", flush=True)
    model = cube(kind, device=DEV, gravity=(0.0, 0.0, -9.81))
    rng = np.random.default_rng(5)
    q = model.host["particle_q"].astype(np.float64) + rng.normal(0, 0.005, (model.particle_count, 3))
    q[:, 2] -= 0.298  # some nodes below the ground plane
    v = rng.normal(0, 0.05, q.shape)
    kw = dict(dt=0.002, ground_plane=(0, 0, 1, 0), ground_ke=10.0, ground_kd=0.02, ground_kf=0.01, ground_mu=0.5,
              linear_damping=0.3)
    mat = (500.0, 1000.0, 0.2)
    model = cube(kind, device=DEV, gravity=(0.0, 0.0, -9.81), material=mat)
    s = solver_for(model, **kw)
    ref = oracle_for(model, **kw)
    st = model.state()
    st.particle_q.copy_(torch.tensor(q, dtype=torch.float32))
    st.particle_qd.copy_(torch.tensor(v, dtype=torch.float32))
    qh, vh = np_(st.particle_q), np_(st.particle_qd)
    term = oracle.TERM_TET if kind == "tet" else oracle.TERM_HEX
    elastic = np_(s.eval_tetrahedral_forces(model, None, st) if kind == "tet" else
                  s.eval_hexahedral_forces(model, None, st))
    assert rel(elastic, ref.eval_terms(term, qh, vh)) < 1e-5
    assert rel(np_(s.eval_particle_ground_contact_forces(model, None, st)),
               ref.eval_terms(oracle.TERM_GROUND, qh, vh)) < 1e-6
    assert rel(np_(s.eval_gravity_forces(model, st)), ref.eval_terms(oracle.TERM_GRAVITY, qh, vh)) < 1e-6
    assert rel(np_(s.eval_linear_damping_forces(model, st)), ref.eval_terms(oracle.TERM_DAMPING, qh, vh)) < 1e-6


@pytest.mark.parametrize("variant", ["resident", "streamed"])
@pytest.mark.parametrize("kind", ["tet", "hex"])
def test_step_matches_oracle_tight_tolerance(kind, variant):
    """Converged solves (tol 1e-7) agree with the FP64 oracle step by step (both CG kernels)."""
    print(f"case step {kind} {variant}", flush=True)
    model = cube(kind, device=DEV, gravity=(0.0, 0.0, -9.81), material=(500.0, 1000.0, 0.1))
    kw = dict(dt=0.002, linear_solver_tolerance=1e-7, linear_solver_maxiter=200)
    s = solver_for(model, **kw)
    s.set_cg_variant(variant)
    assert s.launch_config()["cg_variant"] == variant
    ref = oracle_for(model, **kw)
    rng = np.random.default_rng(11)
    st0, st1 = model.state(), model.state()
    st0.particle_q.add_(torch.tensor(rng.normal(0, 0.002, st0.particle_q.shape), dtype=torch.float32, device=DEV))
    st0.particle_qd.copy_(torch.tensor(rng.normal(0, 0.05, st0.particle_qd.shape), dtype=torch.float32))
    ext = torch.tensor(rng.normal(0, 0.03, st0.particle_qd.shape), dtype=torch.float32, device=DEV)
    q, v = np_(st0.particle_q), np_(st0.particle_qd)
    for k in range(4):
        s.refresh_tangent(st0.particle_q)
        ref.refresh_tangent(q)
        s.step(st0, st1, None, None, 0.002, external_forces=ext)
        out = ref.step(q, v, external_forces=np_(ext))
        assert rel(np_(st0.particle_f), out["f"]) < 1e-5, k
        assert rel(np_(s.dv), ref.dv) < 2e-4, k
        assert rel(np_(st1.particle_qd), out["qd"]) < 2e-4, k
        assert rel(np_(st1.particle_q), out["q"]) < 1e-6, k
        q, v = np_(st1.particle_q), np_(st1.particle_qd)
        st0, st1 = st1, st0


@pytest.mark.parametrize("variant", ["resident", "streamed"])
def test_multichunk_step_matches_oracle(variant):
    """Tet4 drop mesh (1331 nodes) split into 2, 3 and 4 CG chunks: a converged
    step (tol 1e-7) matches the FP64 oracle for every split, and both CG
    kernels take the same iterations at the production tolerance.  Covers
    chunks whose blocks fill the register, LDS and streamed tiers and chunks
    with padded register slots (resident kernel)."""
    print(f"case multichunk {variant}", flush=True)
    sc = meshes.tet_drop_scene()
    model = build_tet_model(sc, device=DEV)
    kw = dict(dt=1 / 600, linear_solver_maxiter=400, linear_solver_tolerance=1e-7, ground_plane=(0, 0, 1, 0),
              ground_ke=100.0, ground_kd=0.1, ground_kf=1.0, ground_mu=0.5)
    ref = oracle_for(model, **kw)
    st0 = model.state()
    rng = np.random.default_rng(21)
    st0.particle_qd.copy_(torch.tensor(rng.normal(0, 0.05, st0.particle_qd.shape), dtype=torch.float32))
    q, v = np_(st0.particle_q), np_(st0.particle_qd)
    ref.refresh_tangent(q)
    out = ref.step(q, v)
    iters = {}
    for wg in (2, 3, 4):
        s = solver_for(model, **kw)
        s.set_launch(0, 0, wg)
        s.set_cg_variant(variant)
        cfg = s.launch_config()
        assert (cfg["cg_variant"], cfg["cg_workgroups"]) == (variant, wg), cfg
        st1 = model.state()
        s.refresh_tangent(st0.particle_q)
        s.step(st0, st1, None, None, 1 / 600)
        err = rel(np_(s.dv), ref.dv)
        print(f"  {wg} chunks: dv rel err {err:.2e}, iterations {s.solve_stats()['iterations'][0]}", flush=True)
        assert err < 2e-4, (wg, err)
        assert rel(np_(st1.particle_q), out["q"]) < 1e-6
        s.linear_solver_tolerance = 1e-2
        s.velocity_increment.zero_()
        s.step(st0, st1, None, None, 1 / 600)
        iters[wg] = s.solve_stats()["iterations"][0]
    other = "streamed" if variant == "resident" else "resident"
    s = solver_for(model, **kw)
    s.set_launch(0, 0, 3)
    s.set_cg_variant(other)
    s.linear_solver_tolerance = 1e-2
    st1 = model.state()
    s.refresh_tangent(st0.particle_q)
    s.step(st0, st1, None, None, 1 / 600)
    assert abs(s.solve_stats()["iterations"][0] - iters[3]) <= 1, (iters, s.solve_stats()["iterations"])


def test_fixed_particles_reduced_dense_solve():
    """Reference test_fixed_particles.py:40-83 on the HIP path."""
    print("case fixed", flush=True)
    model = tet_pair(device=DEV)
    s = solver_for(model, dt=0.002, fixed_particle_indices=(0, 1), linear_solver_maxiter=64)
    cur, nxt = model.state(), model.state()
    rest = model.host["particle_q"].astype(np.float64)
    q = rest.copy()
    q[2] += (0.002, 0.001, -0.001)
    vel = np.array(((0, 0, 0), (0, 0, 0), (0.02, -0.03, 0.01), (-0.01, 0.01, 0.02)), np.float32)
    ext = np.array(((20, 30, 40), (10, -30, 60), (0.02, 0.04, 0.03), (-0.01, 0.03, 0.02)), np.float32)
    forces = torch.tensor(ext, device=DEV)
    cur.particle_q.copy_(torch.tensor(q, dtype=torch.float32))
    cur.particle_qd.copy_(torch.tensor(vel))
    for refresh in (False, True, True):
        if refresh:
            s.refresh_tangent(cur.particle_q)
        A = s.A_bsr.dense()
        elastic = np_(s.eval_tetrahedral_forces(model, None, cur))
        mass = np.repeat(model.host["particle_mass"].astype(np.float64), 3)
        rhs = 0.002 * (elastic + ext).ravel() - (A - np.diag(mass)) @ vel.ravel()
        expected = np.zeros(12)
        expected[6:] = np.linalg.solve(A[6:, 6:], rhs[6:])
        s.step(cur, nxt, None, None, 0.002, external_forces=forces)
        np.testing.assert_allclose(np_(s.velocity_increment).ravel(), expected, rtol=3e-4, atol=2e-6)
        np.testing.assert_array_equal(np_(nxt.particle_q)[:2], rest[:2].astype(np.float32))
        np.testing.assert_array_equal(np_(nxt.particle_qd)[:2], np.zeros((2, 3)))
        np.testing.assert_allclose(np_(nxt.particle_qd), vel + expected.reshape(4, 3), atol=2e-6)
        np.testing.assert_array_equal(s.A_bsr.dense(), A)
        cur, nxt = nxt, cur
        vel = np_(cur.particle_qd).astype(np.float32)


def test_all_fixed_boundary_restored():
    """Reference test_fixed_particles.py:85-94."""
    print("case all-fixed", flush=True)
    model = tet_pair(device=DEV)
    s = solver_for(model, dt=0.002, fixed_particle_indices=(0, 0, 1, 2, 3))
    cur, nxt = model.state(), model.state()
    cur.particle_q.fill_(0.5)
    cur.particle_qd.fill_(100.0)
    s.step(cur, nxt, None, None, 0.002)
    np.testing.assert_array_equal(np_(nxt.particle_q), model.host["particle_q"].astype(np.float64))
    np.testing.assert_array_equal(np_(nxt.particle_qd), np.zeros((4, 3)))
    np.testing.assert_array_equal(np_(s.velocity_increment), np.zeros((4, 3)))


def test_fixed_capture_uses_live_forces_and_exact_anchors():
    """Reference test_fixed_particles.py:92-120: a recorded tangent refresh +
    step replays live external forces against the reduced dense solve of the
    free nodes, with the fixed nodes exactly at their anchors."""
    print("case fixed capture", flush=True)
    model = tet_pair(device=DEV)
    s = solver_for(model, dt=0.002, fixed_particle_indices=(0, 1), linear_solver_maxiter=64)
    cur, nxt = model.state(), model.state()
    forces = torch.zeros(4, 3, dtype=torch.float32, device=DEV)
    s.refresh_tangent(cur.particle_q)
    matrix = s.A_bsr.dense()
    elastic = np_(s.eval_tetrahedral_forces(model, None, cur))
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        s.step(cur, nxt, None, None, 0.002, external_forces=forces)
    torch.cuda.current_stream().wait_stream(side)
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        s.refresh_tangent(cur.particle_q)
        s.step(cur, nxt, None, None, 0.002, external_forces=forces)
    rest = model.host["particle_q"].astype(np.float32)
    for value in (0.02, -0.03, 0.0):
        force = np.zeros((4, 3), dtype=np.float32)
        force[2, 2] = value
        forces.copy_(torch.tensor(force))
        graph.replay()
        torch.cuda.synchronize()
        np.testing.assert_array_equal(np_(nxt.particle_q)[:2], rest[:2])
        np.testing.assert_array_equal(np_(nxt.particle_qd)[:2], np.zeros((2, 3)))
        expected = np.zeros(12)
        rhs = 0.002 * (elastic + force).ravel()
        expected[6:] = np.linalg.solve(matrix[6:, 6:], rhs[6:])
        np.testing.assert_allclose(np_(nxt.particle_qd).ravel(), expected, rtol=5e-4, atol=3e-8)
        if value:
            assert np_(nxt.particle_qd)[2, 2] * value > 0


def test_true_residual_within_tolerance():
    """||A dv - b|| / ||b|| <= tol for the production 1e-2 solve (test_solver.py:403-405)."""
    print("case residual", flush=True)
    model = cube("tet", device=DEV, gravity=(0, 0, -9.81), material=(500.0, 1000.0, 0.1))
    s = solver_for(model, dt=0.002, linear_solver_maxiter=128)
    s.linear_solver_check_every = 1
    st0, st1 = model.state(), model.state()
    rng = np.random.default_rng(13)
    ext = torch.tensor(rng.normal(0, 0.03, st0.particle_q.shape), dtype=torch.float32, device=DEV)
    for tol in (1e-2, 1e-5):
        s.linear_solver_tolerance = tol
        s.velocity_increment.zero_()
        st0, st1 = model.state(), model.state()
        s.step(st0, st1, None, None, 0.002, external_forces=ext)
        A = s.A_bsr.dense()
        rhs = np_(st0.particle_f).ravel()
        res = np.linalg.norm(A @ np_(s.dv).ravel() - rhs) / np.linalg.norm(rhs)
        assert res <= tol * 1.01, (tol, res, s.last_cg_iters)


def test_iteration_count_semantics():
    """linear_solver_check_every > 0 (solver.py:1203-1217, 1333-1335): early exit, and
    last_cg_iters is the first multiple of check_every at or after convergence; with
    check_every = 0 and no CUDA-graph exit the full budget runs (residual-masked),
    giving the same increment."""
    print("case iteration counts", flush=True)
    model = cube("tet", device=DEV, gravity=(0, 0, -9.81), material=(500.0, 1000.0, 0.1))
    rng = np.random.default_rng(14)
    ext = torch.tensor(rng.normal(0, 0.03, (model.particle_count, 3)), dtype=torch.float32, device=DEV)
    out = {}
    for every, graph in ((1, False), (4, False), (0, True), (0, False)):
        s = solver_for(model, dt=0.002, linear_solver_maxiter=64, linear_solver_tolerance=1e-5,
                       linear_solver_use_cuda_graph=graph)
        s.linear_solver_check_every = every
        st0, st1 = model.state(), model.state()
        s.step(st0, st1, None, None, 0.002, external_forces=ext)
        out[(every, graph)] = (s.last_cg_iters, s.solve_stats()["iterations"][0], np_(s.dv))
    exact = out[(1, False)][0]
    assert 0 < exact < 64 and out[(1, False)][1] == exact
    assert out[(4, False)][0] == -(-exact // 4) * 4
    assert out[(0, True)][1] == exact and out[(0, True)][0] is None
    assert out[(0, False)][1] == exact  # converged updates masked: counts the live iterations
    for key in out:
        np.testing.assert_array_equal(out[key][2], out[(1, False)][2])


def test_drop_rollout_positions():
    """Tet4 drop (10^3 cells x 5 tets): 60 frames x 10 substeps vs the FP64 oracle."""
    print("case drop rollout", flush=True)
    sc = meshes.tet_drop_scene()
    model = build_tet_model(sc, device=DEV)
    kw = dict(dt=1 / 600, linear_solver_maxiter=50, linear_solver_tolerance=1e-2, ground_plane=(0, 0, 1, 0),
              ground_ke=100.0, ground_kd=0.1, ground_kf=1.0, ground_mu=0.5)
    s = solver_for(model, **kw)
    ref = oracle_for(model, **kw)
    st0, st1 = model.state(), model.state()
    q, v = np_(st0.particle_q), np_(st0.particle_qd)
    worst = 0.0
    for k in range(600):
        s.refresh_tangent(st0.particle_q)
        s.step(st0, st1, None, None, 1 / 600)
        ref.refresh_tangent(q)
        out = ref.step(q, v)
        q, v = out["q"], out["qd"]
        st0, st1 = st1, st0
        if k % 50 == 49:
            worst = max(worst, rel(np_(st0.particle_q), q))
    print(f"drop rollout worst position rel err {worst:.2e}", flush=True)
    assert worst < 1e-4
    assert np.isfinite(np_(st0.particle_q)).all()


def test_hex_cantilever_rollout():
    """Hex8 64x16x16 cantilever, x=0 fixed: 100 substeps vs the FP64 oracle (positions 1e-4 relative)."""
    print("case cantilever", flush=True)
    sc = meshes.hex_cantilever_scene()
    model = build_hex_model(sc, device=DEV)
    kw = dict(dt=1 / 600, linear_solver_maxiter=50, linear_solver_tolerance=1e-2)
    s = solver_for(model, fixed_particle_indices=sc.fixed, **kw)
    ref = oracle_for(model, fixed=sc.fixed, **kw)
    ref.set_threads(16)
    st0, st1 = model.state(), model.state()
    q, v = np_(st0.particle_q), np_(st0.particle_qd)
    worst = 0.0
    for k in range(100):
        s.refresh_tangent(st0.particle_q)
        s.step(st0, st1, None, None, 1 / 600)
        ref.refresh_tangent(q)
        out = ref.step(q, v)
        q, v = out["q"], out["qd"]
        st0, st1 = st1, st0
        worst = max(worst, rel(np_(st0.particle_q), q))
        if k % 25 == 24:
            disp = np.abs(np_(st0.particle_q) - model.host["particle_q"]).max()
            print(f"  substep {k + 1}: worst position rel err {worst:.2e}, tip displacement {disp:.3e}", flush=True)
    assert worst < 1e-4
    np.testing.assert_array_equal(np_(st0.particle_q)[sc.fixed], model.host["particle_q"][sc.fixed])


def _large_step(s, model, dt, tol, maxiter, early_exit):
    # Device-side residual exit follows linear_solver_use_cuda_graph (reference
    # solver.py:1205-1216); the setters below re-send the parameters.
    s.linear_solver_use_cuda_graph = early_exit
    s.linear_solver_tolerance = tol
    s.linear_solver_maxiter = maxiter
    st0, st1 = model.state(), model.state()
    s.velocity_increment.zero_()
    s.refresh_tangent(st0.particle_q)
    s.step(st0, st1, None, None, dt)
    return st0, st1


_LARGE_ORACLE = {}


@pytest.mark.parametrize("variant", ["resident", "streamed"])
@pytest.mark.parametrize("early_exit", [True, False])
def test_large_substep_matches_oracle(early_exit, variant):
    """1M-tet solid (x=0 face fixed): one substep on every XCD vs the FP64 oracle.

    A converged solve (tol 1e-6) must reproduce the oracle's velocity increment;
    the fixed 50-iteration budget (residual exit disabled) must agree with the
    oracle's 50 iterations.  Either fails loudly if a workgroup reads a stale
    exchange record or partial sum, or (resident kernel) a block held on chip
    is attributed to the wrong row.
    """
    print(f"case large substep early_exit={early_exit} {variant}", flush=True)
    sc = meshes.large_tet_scene()
    model = build_tet_model(sc, device=DEV)
    dt = 1 / 600
    s = solver_for(model, dt=dt, fixed_particle_indices=sc.fixed, linear_solver_use_cuda_graph=early_exit)
    s.set_cg_variant(variant)
    assert s.launch_config()["cg_variant"] == variant
    tol, maxiter = (1e-6, 400) if early_exit else (1e-12, 50)
    st0, st1 = _large_step(s, model, dt, tol, maxiter, early_exit)
    if early_exit not in _LARGE_ORACLE:
        ref = oracle_for(model, fixed=sc.fixed, dt=dt, linear_solver_tolerance=tol, linear_solver_maxiter=maxiter)
        ref.refresh_tangent(np_(st0.particle_q))
        out = ref.step(np_(st0.particle_q), np_(st0.particle_qd))
        _LARGE_ORACLE[early_exit] = (out, ref.dv.copy())
    out, ref_dv = _LARGE_ORACLE[early_exit]
    err_dv = rel(np_(s.dv), ref_dv)
    print(f"large substep dv rel err {err_dv:.2e} (oracle iterations {out['iters']}, {s.launch_config()})", flush=True)
    assert err_dv < (2e-3 if early_exit else 1e-3)
    assert rel(np_(st1.particle_q), out["q"]) < 1e-6
    np.testing.assert_array_equal(np_(st1.particle_q)[sc.fixed], model.host["particle_q"][sc.fixed])


def test_cg_launch_shapes_agree():
    """The persistent CG's result does not depend on how rows are split into workgroups."""
    print("case launch shapes", flush=True)
    sc = meshes.large_tet_scene()
    model = build_tet_model(sc, device=DEV)
    dt = 1 / 600
    s = solver_for(model, dt=dt, fixed_particle_indices=sc.fixed)
    base = None
    for variant, wg in (("streamed", 128), ("streamed", 256), ("streamed", 416), ("streamed", 512),
                        ("resident", 256)):
        s.set_launch(0, 0, wg)
        s.set_cg_variant(variant)
        cfg = s.launch_config()
        assert cfg["cg_variant"] == variant, cfg
        _large_step(s, model, dt, 1e-6, 400, True)
        dv = np_(s.dv)
        if base is None:
            base = dv
        else:
            assert rel(dv, base) < 1e-3, (variant, wg)
