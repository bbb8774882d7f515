"""Drop-in boundary details on the HIP path.

* ``M_bsr.matvec(x, y, z, alpha, beta)`` is the reference's block-Jacobi
  operator (solver.py:1175-1201, block_jacobi_mv_kernel kernels.py:1852-1868):
  z = alpha M^-1 x + beta y, and with alpha 1, beta 0 it equals the prepared
  step's apply_preconditioner bit for bit (reference
  tests/test_fem_prepared_step.py:124-126).
* The default ``self_contact_ke = 1e3`` on an FEM mesh with particle radii
  skips particle-particle contact as the reference does for models without a
  hash grid (solver.py:1477-1485): the step equals one with the term off.
* A persistent CG launch whose waits run out (forced from the host with a
  one-poll bound on a four-workgroup solve; every kernel still exits
  normally) leaves NaN in the output state instead of a partial solve, and
  solve_stats() raises; the next solve after a reset is exact again.
* Error behaviour of the reference's step and tangent refresh: a step of a
  non-differentiable solver on inputs that record gradients is refused
  (solver.py:985-986), and so is a timestep change while a CUDA graph is
  being captured (solver.py:801-802); nothing is recorded by either.
"""

import numpy as np
import pytest
import torch

from helpers import cube
from newton_solver_fem_amd import build_tet_model, meshes
from newton_solver_fem_amd._lib import FEMError

pytestmark = pytest.mark.gpu

DEV = "cuda:0"


def np_(t):
    return t.detach().cpu().numpy().astype(np.float64)


def solver_for(model, **kw):
    from newton_solver_fem_amd.solver import SolverFEM

    base = dict(solver_type="cg", preconditioner_type="block_diag")
    base.update(kw)
    return SolverFEM(model, **base)


def test_block_jacobi_operator_matvec():
    print("case M_bsr.matvec", flush=True)
    model = cube("tet", device=DEV, material=(500.0, 1000.0, 0.1), gravity=(0.0, 0.0, -9.81))
    s = solver_for(model, dt=0.002, self_contact_ke=0.0)
    rng = np.random.default_rng(3)
    q = model.particle_q + torch.tensor(rng.normal(0, 0.003, (model.particle_count, 3)), dtype=torch.float32,
                                        device=DEV)
    s.refresh_tangent(q)
    n = model.particle_count
    x = torch.tensor(rng.normal(size=(n, 3)), dtype=torch.float32, device=DEV)
    y = torch.tensor(rng.normal(size=(n, 3)), dtype=torch.float32, device=DEV)
    z = torch.empty_like(x)
    minv = np_(s._pc_inv_diag)
    exp = np.einsum("kij,kj->ki", minv, np_(x))
    s.M_bsr.matvec(x, y, z, 0.7, -1.3)
    np.testing.assert_allclose(np_(z), 0.7 * exp - 1.3 * np_(y), rtol=1e-5, atol=1e-6 * np.abs(exp).max())
    s.M_bsr.matvec(x, y, z, 1.0, 0.0)
    st = model.state()
    st.particle_q.copy_(q)
    prepared = s.prepare_step(st, None, None, 0.002)
    out = torch.empty_like(x)
    prepared.apply_preconditioner(x, out)
    np.testing.assert_array_equal(np_(z), np_(out))
    y0 = np_(y)
    s.M_bsr.matvec(x, y, y, 2.0, 1.0)  # in place on y
    np.testing.assert_allclose(np_(y), 2.0 * np_(z) + y0, rtol=1e-6, atol=1e-7)
    with pytest.raises(ValueError):
        s.M_bsr.matvec(x[:-1], y, z, 1.0, 0.0)


def test_default_self_contact_is_skipped_for_fem_meshes():
    print("case default self-contact", flush=True)
    model = cube("tet", device=DEV, material=(500.0, 1000.0, 0.1), gravity=(0.0, 0.0, -9.81))
    assert float(model.particle_radius.max()) > 0 and model.particle_grid is None
    a = solver_for(model, dt=0.002)  # reference default self_contact_ke = 1e3
    b = solver_for(model, dt=0.002, self_contact_ke=0.0)
    assert model.particle_ke == 0.0 and model.particle_kd == 100.0
    outs = []
    for s in (a, b):
        s0, s1 = model.state(), model.state()
        s0.particle_qd[:, 2] = -0.2
        s.refresh_tangent(s0.particle_q)
        s.step(s0, s1, None, None, 0.002)
        outs.append((np_(s1.particle_q), np_(s1.particle_qd)))
    np.testing.assert_array_equal(outs[0][0], outs[1][0])
    np.testing.assert_array_equal(outs[0][1], outs[1][1])


@pytest.mark.parametrize("variant", ["resident", "streamed"])
def test_barrier_timeout_leaves_nan_not_a_plausible_state(variant):
    print(f"case barrier timeout {variant}", flush=True)
    sc = meshes.tet_drop_scene()
    model = build_tet_model(sc, device=DEV)
    kw = dict(dt=1 / 600, self_contact_ke=0.0, linear_solver_tolerance=1e-6, linear_solver_maxiter=100,
              ground_plane=(0, 0, 1, 0), ground_ke=100.0, ground_kd=0.1, ground_kf=1.0, ground_mu=0.5)
    s = solver_for(model, **kw)
    s.set_launch(0, 0, 4)
    s.set_cg_variant(variant)
    assert s.launch_config()["cg_workgroups"] == 4
    s0, s1 = model.state(), model.state()
    s0.particle_qd[:, 2] = -0.3
    s.refresh_tangent(s0.particle_q)
    s.set_barrier_spin_limit(1)
    s.step(s0, s1, None, None, 1 / 600)
    q = np_(s1.particle_q)
    print(f"  NaN rows after the forced timeout: {int(np.isnan(q).any(axis=1).sum())} of {len(q)}", flush=True)
    assert np.isnan(q).any() and np.isnan(np_(s.dv)).any()
    with pytest.raises(FEMError, match="barrier"):
        s.solve_stats()
    s.set_barrier_spin_limit(0)
    s.velocity_increment.zero_()
    s.step(s0, s1, None, None, 1 / 600)
    fresh = solver_for(model, **kw)
    fresh.set_launch(0, 0, 4)
    fresh.set_cg_variant(variant)
    f1 = model.state()
    fresh.refresh_tangent(s0.particle_q)
    fresh.step(s0, f1, None, None, 1 / 600)
    assert s.solve_stats()["iterations"][0] > 0
    np.testing.assert_array_equal(np_(s1.particle_q), np_(f1.particle_q))
    np.testing.assert_array_equal(np_(s1.particle_qd), np_(f1.particle_qd))


def test_step_refuses_gradient_inputs_and_capture_refuses_a_timestep_change(monkeypatch):
    print("case boundary tape and capture guards", flush=True)
    model = cube("tet", density=100.0, gravity=(0.0, 0.0, -9.81), device=DEV)
    s = solver_for(model, dt=0.002, self_contact_ke=0.0)
    a, b = model.state(), model.state()
    s.refresh_tangent(a.particle_q)
    a.particle_q.requires_grad_(True)
    with pytest.raises(RuntimeError, match="differentiable=True before recording"):
        s.step(a, b, None, None, 0.002)
    with torch.no_grad():  # outside autograd recording the step runs
        s.step(a, b, None, None, 0.002)
    a.particle_q.requires_grad_(False)
    # the capture state is simulated: nothing is recorded
    monkeypatch.setattr(torch.cuda, "is_current_stream_capturing", lambda: True)
    with pytest.raises(RuntimeError, match="outside tapes and CUDA graph capture"):
        s.refresh_tangent(a.particle_q, 0.001)
