"""Batched independent meshes in one solver (block-diagonal BSR, per-instance CG).

A batch packs uneven Tet4 meshes back to back (``instance_offsets``); every
instance runs its own PCG (own dots, alpha/beta, residual exit and
statistics) inside one persistent launch; when the chunks outnumber the
workgroups that fit at once, the workgroups take whole instances' chunks from
a device queue.  With the
same chunking per instance, the batch reproduces the meshes stepped one by
one (looped) bit for bit.
"""

import numpy as np
import pytest
import torch

from newton_solver_fem_amd import build_tet_model, merge_tet_scenes, meshes

pytestmark = pytest.mark.gpu

DEV = "cuda:0"
GROUND = dict(ground_plane=(0.0, 0.0, 1.0, 0.0), ground_ke=100.0, ground_kd=0.1, ground_kf=1.0, ground_mu=0.5)


def solver_for(model, **kw):
    from newton_solver_fem_amd.solver import SolverFEM

    base = dict(solver_type="cg", preconditioner_type="block_diag", self_contact_ke=0.0, **GROUND)
    base.update(kw)
    return SolverFEM(model, **base)


def rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30))


def np_(t):
    return t.detach().cpu().numpy().astype(np.float64)


def rollout(s, model, steps, dt):
    st0, st1 = model.state(), model.state()
    for _ in range(steps):
        s.refresh_tangent(st0.particle_q)
        s.step(st0, st1, None, None, dt)
        st0, st1 = st1, st0
    return st0


@pytest.fixture(scope="module")
def batch():
    scenes = meshes.batched_tet_scenes(count=16, seed=5, t_range=(300, 6000))
    merged, offs = merge_tet_scenes(scenes)
    return scenes, merged, offs


@pytest.mark.parametrize("variant", ["streamed", "resident"])
def test_batched_bit_identical_to_looped(batch, variant):
    """Same chunks per instance on both sides: per-instance arithmetic is identical, so is the result."""
    scenes, merged, offs = batch
    print(f"case batched == looped {variant}", flush=True)
    dt, steps = 1 / 600, 4
    kw = dict(dt=dt, linear_solver_tolerance=1e-2, linear_solver_maxiter=50)
    model = build_tet_model(merged, device=DEV)
    s = solver_for(model, instance_offsets=offs, **kw)
    s.set_cg_variant(variant)
    s.set_launch(0, 0, len(scenes))  # the minimum: every instance gets the chunks it needs, no more
    assert s.launch_config()["cg_variant"] == variant
    out = rollout(s, model, steps, dt)
    it_b = s.solve_stats()["iterations"]
    assert len(it_b) == len(scenes) and min(it_b) > 0
    dv_b, q_b = np_(s.dv), np_(out.particle_q)
    looped_chunks = 0
    for k, sc in enumerate(scenes):
        m = build_tet_model(sc, device=DEV)
        sk = solver_for(m, **kw)
        sk.set_cg_variant(variant)
        sk.set_launch(0, 0, 1)
        looped_chunks += sk.launch_config()["cg_workgroups"]
        ok = rollout(sk, m, steps, dt)
        lo, hi = int(offs[k]), int(offs[k + 1])
        np.testing.assert_array_equal(dv_b[lo:hi], np_(sk.dv), err_msg=f"instance {k}")
        np.testing.assert_array_equal(q_b[lo:hi], np_(ok.particle_q), err_msg=f"instance {k}")
        assert sk.solve_stats()["iterations"][0] == it_b[k]
    assert s.launch_config()["cg_workgroups"] == looped_chunks
    if variant == "resident":  # both kernels agree on every instance's iteration count
        t = solver_for(model, instance_offsets=offs, **kw)
        t.set_cg_variant("streamed")
        t.set_launch(0, 0, len(scenes))
        rollout(t, model, steps, dt)
        np.testing.assert_array_equal(np.asarray(t.solve_stats()["iterations"]), np.asarray(it_b))


def test_batched_chunk_queue(batch):
    """More chunks than workgroups fit at once: one launch whose workgroups take chunks from a device queue.

    Chunking differs from the default launch, so the fp32 reduction order
    differs; both solves are converged to tol 1e-6, and two solutions with
    relative residual <= tol differ by at most ~2 tol cond(M^-1 A) (< 1e-3 here).
    Which workgroup takes which chunk changes from run to run, the chunks'
    arithmetic does not: two queued solvers agree bit for bit.
    """
    scenes, merged, offs = batch
    print("case chunk queue", flush=True)
    dt = 1 / 600
    kw = dict(dt=dt, linear_solver_tolerance=1e-6, linear_solver_maxiter=400, instance_offsets=offs)
    model = build_tet_model(merged, device=DEV)
    a = solver_for(model, **kw)
    b = solver_for(model, **kw)
    c = solver_for(model, **kw)
    for s in (b, c):
        s.set_launch(0, 0, 1100)
        cfg = s.launch_config()
        assert cfg["cg_launch_groups"] == 1 and cfg["cg_queued"]
        assert cfg["cg_launch_workgroups"] < cfg["cg_workgroups"]
    assert not a.launch_config()["cg_queued"]
    qa, qb = np_(rollout(a, model, 2, dt).particle_q), np_(rollout(b, model, 2, dt).particle_q)
    qc = np_(rollout(c, model, 2, dt).particle_q)
    assert rel(np_(b.dv), np_(a.dv)) < 1e-3
    assert rel(qb, qa) < 1e-6
    np.testing.assert_array_equal(np_(c.dv), np_(b.dv))
    np.testing.assert_array_equal(qc, qb)
    sa, sb = a.solve_stats(), b.solve_stats()
    assert len(sb["iterations"]) == len(scenes)
    assert sb["iterations"] == c.solve_stats()["iterations"]
    assert (np.asarray(sb["residual_sq"]) <= np.asarray(sb["atol_sq"]) * 1.0001).all()


def test_per_instance_residual_exit(batch):
    """Each instance stops at its own tolerance: iteration counts differ per instance."""
    scenes, merged, offs = batch
    print("case per-instance exit", flush=True)
    model = build_tet_model(merged, device=DEV)
    s = solver_for(model, instance_offsets=offs, dt=1 / 600, linear_solver_tolerance=1e-4,
                   linear_solver_maxiter=200)
    rollout(s, model, 3, 1 / 600)
    stats = s.solve_stats()
    it = np.asarray(stats["iterations"])
    assert len(set(it.tolist())) > 1, it
    rr, atol = np.asarray(stats["residual_sq"]), np.asarray(stats["atol_sq"])
    assert (rr <= atol * 1.0001).all(), (rr, atol)


def test_autotune_keeps_results(batch):
    """The autotuner only changes launch shapes: results before and after agree."""
    scenes, merged, offs = batch
    print("case autotune", flush=True)
    model = build_tet_model(merged, device=DEV)
    kw = dict(dt=1 / 600, linear_solver_tolerance=1e-6, linear_solver_maxiter=400, instance_offsets=offs)
    s = solver_for(model, **kw)
    ref = np_(rollout(s, model, 3, 1 / 600).particle_q)
    s.reset()
    cfg = s.autotune(model.state(), 1 / 600, repeats=1)
    assert cfg["cg_workgroups"] >= len(scenes)
    got = np_(rollout(s, model, 3, 1 / 600).particle_q)
    assert rel(got, ref) < 1e-6


def test_graph_replay_matches_eager():
    """A HIP graph of two substeps replays to the eager result, every replay."""
    print("case graph replay", flush=True)
    sc = meshes.tet_drop_scene()
    dt = 1 / 600
    kw = dict(dt=dt, linear_solver_tolerance=1e-2, linear_solver_maxiter=50)
    model = build_tet_model(sc, device=DEV)
    eager = solver_for(model, **kw)
    ref = []
    st0, st1 = model.state(), model.state()
    for k in range(8):
        eager.refresh_tangent(st0.particle_q, dt)
        eager.step(st0, st1, None, None, dt)
        st0, st1 = st1, st0
        if k % 2 == 1:
            ref.append(np_(st0.particle_q))
    s = solver_for(model, **kw)
    s0, s1 = model.state(), model.state()

    def body():
        s.refresh_tangent(s0.particle_q, dt)
        s.step(s0, s1, None, None, dt)
        s.refresh_tangent(s1.particle_q, dt)
        s.step(s1, s0, None, None, dt)

    # Capture from the rest state: the warm-up pass runs on copies.
    w0, w1 = model.state(), model.state()
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        s.refresh_tangent(w0.particle_q, dt)
        s.step(w0, w1, None, None, dt)
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    s.reset()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        body()
    for k in range(4):
        g.replay()
        torch.cuda.synchronize()
        got = np_(s0.particle_q)
        np.testing.assert_array_equal(got, ref[k])
