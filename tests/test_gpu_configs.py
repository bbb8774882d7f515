"""The benchmark workloads at their full sizes on the HIP path, against the FP64 oracle.

* Large Tet4 solid (60x60x56 cells x 5 tets, 1,008,000 tets, jittered,
  log-uniform E, x=0 face fixed): 10 substeps of tangent refresh + step
  (tol 1e-2, <= 50 PCG iterations) vs the oracle's, positions within 1e-4
  relative after every substep.
* Batched uneven meshes (256 Tet4 lattices, tets log-uniform in
  [500, 50,000], seed 1) in one solver, 2 substeps at tol 1e-2: every
  instance's own true residual <= tol, the batch equals the 256 meshes
  stepped one by one bit for bit, and 8 instances spanning the size range
  (the largest included) match the oracle's positions within 1e-4.
* Differentiable bar (40x8x8 cells, 12,800 tets, x=0 face fixed, seeded
  activations scaling mu, 60 substeps, checkpoint every 10): d(tip-target
  loss)/d(activations) in 3 random directions and d loss/d mu in one
  direction vs FP64 central differences (eps 1e-4) of the oracle rollout,
  1e-3 relative (reference examples/diffsim/_bear_training.py:254-310,
  differentiation.py:287-394).

The Hex8 cantilever at its full size and 100 substeps is in
test_gpu_parity.py::test_hex_cantilever_rollout.
"""

import os

import numpy as np
import pytest
import torch

import oracle
from newton_solver_fem_amd import build_tet_model, merge_tet_scenes, meshes

pytestmark = pytest.mark.gpu

DEV = "cuda:0"
DT = 1.0 / 600.0
GROUND = dict(ground_plane=(0.0, 0.0, 1.0, 0.0), ground_ke=100.0, ground_kd=0.1, ground_kf=1.0, ground_mu=0.5)
# Host threads of the oracle's loops (the checker only; results differ from
# one thread by FP64 summation order).
ORACLE_THREADS = min(16, os.cpu_count() or 1)


def rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30))


def np_(t):
    return t.detach().cpu().numpy().astype(np.float64)


def solver_for(model, **kw):
    from newton_solver_fem_amd.solver import SolverFEM

    base = dict(solver_type="cg", preconditioner_type="block_diag", self_contact_ke=0.0)
    base.update(kw)
    return SolverFEM(model, **base)


def oracle_for(host, gravity, fixed=(), tol=1e-2, maxiter=50, ground=None):
    g = ground or {}
    o = oracle.OracleSolver(host, bits=64, dt=DT, gravity=gravity, fixed=fixed, tolerance=tol, maxiter=maxiter, **g)
    o.set_threads(ORACLE_THREADS)
    return o


def bsr_residual(A, x, b, offs):
    """Per-instance ||A x - b|| / ||b|| with the BSR view (FP64 on the host)."""
    off = A.offsets.cpu().numpy().astype(np.int64)
    cols = A.columns.cpu().numpy().astype(np.int64)
    vals = A.values.detach().cpu().numpy().astype(np.float64)
    rows = np.repeat(np.arange(len(off) - 1), np.diff(off))
    prod = np.einsum("kij,kj->ki", vals, x[cols])
    ax = np.zeros_like(x)
    np.add.at(ax, rows, prod)
    r = ax - b
    out = []
    for k in range(len(offs) - 1):
        lo, hi = int(offs[k]), int(offs[k + 1])
        out.append(np.linalg.norm(r[lo:hi]) / max(np.linalg.norm(b[lo:hi]), 1e-300))
    return np.array(out)


def test_large_tet_solid_ten_substeps_match_oracle():
    print("case large solid 10 substeps", flush=True)
    sc = meshes.large_tet_scene()
    model = build_tet_model(sc, device=DEV)
    s = solver_for(model, dt=DT, fixed_particle_indices=sc.fixed, linear_solver_maxiter=50,
                   linear_solver_tolerance=1e-2)
    o = oracle_for(model.host, model.gravity, fixed=sc.fixed)
    st0, st1 = model.state(), model.state()
    q, v = np_(st0.particle_q), np_(st0.particle_qd)
    worst = 0.0
    for k in range(10):
        s.refresh_tangent(st0.particle_q)
        s.step(st0, st1, None, None, DT)
        o.refresh_tangent(q)
        out = o.step(q, v)
        q, v = out["q"], out["qd"]
        st0, st1 = st1, st0
        err = rel(np_(st0.particle_q), q)
        disp = rel(np_(st0.particle_q) - model.host["particle_q"], q - model.host["particle_q"])
        it = s.solve_stats()["iterations"][0]
        print(f"  substep {k}: position rel err {err:.2e} (displacement {disp:.2e}), iterations {it} / "
              f"oracle {out['iters']}", flush=True)
        worst = max(worst, err)
    assert worst < 1e-4
    assert rel(np_(st0.particle_qd), v) < 2e-2  # tol-1e-2 solves: velocities agree to the tolerance
    np.testing.assert_array_equal(np_(st0.particle_q)[sc.fixed], model.host["particle_q"][sc.fixed])


def _steps(s, model, n, state=None):
    st0, st1 = model.state(), model.state()
    for _ in range(n):
        s.refresh_tangent(st0.particle_q)
        s.step(st0, st1, None, None, DT)
        st0, st1 = st1, st0
    return st1, st0  # (input of the last step, its output)


def test_batched_256_uneven_meshes():
    print("case batched 256 meshes", flush=True)
    scenes = meshes.batched_tet_scenes()
    merged, offs = merge_tet_scenes(scenes)
    assert len(scenes) == 256
    kw = dict(dt=DT, linear_solver_tolerance=1e-2, linear_solver_maxiter=50, **GROUND)
    model = build_tet_model(merged, device=DEV)
    s = solver_for(model, instance_offsets=offs, **kw)
    s.set_launch(0, 0, len(scenes))  # the fewest chunks per instance, as the looped solvers below get
    last_in, out = _steps(s, model, 2)
    stats = s.solve_stats()
    it = np.asarray(stats["iterations"])
    assert len(it) == 256 and it.min() > 0 and it.max() <= 50
    res = bsr_residual(s.A_bsr, np_(s.dv), np_(last_in.particle_f), offs)
    print(f"  iterations {it.min()}..{it.max()} (mean {it.mean():.1f}), worst true residual {res.max():.3e}, "
          f"launch {s.launch_config()}", flush=True)
    assert res.max() <= 1e-2 * 1.01
    dv_b, q_b, qd_b = np_(s.dv), np_(out.particle_q), np_(out.particle_qd)
    cfg = s.launch_config()

    for k, sc in enumerate(scenes):
        m = build_tet_model(sc, device=DEV)
        sk = solver_for(m, **kw)
        # Same lanes per row in the force kernel (their count sets the order
        # of a row's force sums) and the fewest CG chunks.
        sk.set_launch(cfg["elem_lanes"], cfg["assemble_rows"], 1)
        _, ok = _steps(sk, m, 2)
        lo, hi = int(offs[k]), int(offs[k + 1])
        np.testing.assert_array_equal(dv_b[lo:hi], np_(sk.dv), err_msg=f"instance {k}")
        np.testing.assert_array_equal(q_b[lo:hi], np_(ok.particle_q), err_msg=f"instance {k}")
        assert sk.solve_stats()["iterations"][0] == it[k], k
        del sk, m

    sizes = np.array([len(sc.tets) for sc in scenes])
    order = np.argsort(sizes)
    pick = sorted(set(order[np.linspace(0, 255, 8).astype(int)].tolist()) | {int(np.argmax(sizes))})
    assert len(pick) >= 8
    for k in pick:
        sc = scenes[k]
        host = build_tet_model(sc, device="cpu").host
        o = oracle_for(host, model.gravity, ground=GROUND)
        q = host["particle_q"].astype(np.float64)
        v = np.zeros_like(q)
        for _ in range(2):
            o.refresh_tangent(q)
            r = o.step(q, v)
            q, v = r["q"], r["qd"]
        lo, hi = int(offs[k]), int(offs[k + 1])
        e_q = rel(q_b[lo:hi], q)
        e_v = rel(qd_b[lo:hi], v)
        print(f"  instance {k} ({sizes[k]} tets): position rel err {e_q:.2e}, velocity {e_v:.2e}, "
              f"iterations {it[k]} / oracle {r['iters']}", flush=True)
        assert e_q < 1e-4, k
        assert e_v < 2e-2, k


def test_differentiable_bar_gradient_matches_finite_differences():
    print("case differentiable bar 40x8x8, 60 substeps", flush=True)
    from newton_solver_fem_amd.differentiation import rollout

    sc = meshes.tet_bar_scene()
    model = build_tet_model(sc, device=DEV)
    assert model.tet_count == 12_800 and len(sc.fixed) == 81
    steps, every = 60, 10
    s = solver_for(model, dt=DT, differentiable=True, defer_differentiation_checks=True, linear_solver_maxiter=256,
                   fixed_particle_indices=sc.fixed, allow_fixed_in_differentiation=True)
    rest = model.host["particle_q"].astype(np.float64)
    tip = rest[:, 0] > rest[:, 0].max() - 1e-9
    tgt = rest[tip].mean(0) + np.random.default_rng(3).uniform(-0.01, 0.01, 3)
    mat0 = torch.tensor(sc.materials, dtype=torch.float32, device=DEV)
    mu = mat0[:, 0].clone().requires_grad_(True)
    a = torch.tensor(sc.activations, dtype=torch.float32, device=DEV, requires_grad=True)
    s.set_tet_materials(torch.stack([mu * (1.0 + a), mat0[:, 1], mat0[:, 2]], dim=1))
    rest_t = model.particle_q.clone()
    qT, _ = rollout(s, rest_t.clone(), torch.zeros_like(rest_t), steps, checkpoint_every=every)
    loss = ((qT[torch.tensor(tip, device=DEV)].mean(0) - torch.tensor(tgt, dtype=torch.float32, device=DEV)) ** 2
            ).sum()
    loss.backward()
    res = s.check_differentiation()
    print(f"  loss {loss.item():.6e}, residuals {res}", flush=True)

    o = oracle.OracleSolver(model.host, bits=64, dt=DT, gravity=model.gravity, fixed=sc.fixed)
    o.set_threads(ORACLE_THREADS)
    mu0 = sc.materials[:, 0].astype(np.float64)

    def oracle_loss(acts, mus=mu0):
        m = sc.materials.astype(np.float64).copy()
        m[:, 0] = mus * (1.0 + acts)
        o.set_tet_materials(m)
        o.refresh_tangent(rest)
        q, v = rest.copy(), np.zeros_like(rest)
        for _ in range(steps):
            q, v, _ = o.diff_step(q, v)
        return float(np.sum((q[tip].mean(0) - tgt) ** 2))

    a0 = sc.activations.astype(np.float64)
    base = oracle_loss(a0)
    assert abs(loss.item() - base) <= 1e-4 * base, (loss.item(), base)
    rng = np.random.default_rng(4)
    eps = 1e-4
    worst = 0.0
    for k in range(3):
        d = rng.normal(size=model.tet_count)
        fd = (oracle_loss(a0 + eps * d) - oracle_loss(a0 - eps * d)) / (2 * eps)
        exact = float(np.sum(np_(a.grad) * d))
        err = abs(exact - fd) / abs(fd)
        print(f"  d loss / d act direction {k}: adjoint {exact:.6e}, FD {fd:.6e}, rel err {err:.2e}", flush=True)
        worst = max(worst, err)
    d = rng.normal(size=model.tet_count) * mu0
    fd = (oracle_loss(a0, mu0 + eps * d) - oracle_loss(a0, mu0 - eps * d)) / (2 * eps)
    exact = float(np.sum(np_(mu.grad) * d))
    err = abs(exact - fd) / abs(fd)
    print(f"  d loss / d mu: adjoint {exact:.6e}, FD {fd:.6e}, rel err {err:.2e}", flush=True)
    assert worst < 1e-3 and err < 1e-3
