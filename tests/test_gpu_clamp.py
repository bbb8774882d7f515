"""The stretch clamp while it binds, on the HIP path vs the FP64 oracle.

`clamp_deformation_stretch` (reference kernels.py:421-444) replaces the
deformation gradient by U diag(clamp(sigma, min, max)) V^T (signed SVD: an
inverted element has a negative smallest singular value), and its adjoint
(kernels.py:447-481) is the spectral formula.  Every case below has
singular values outside the bounds, so the device SVD slow path, the
fast-path test that skips it and the spectral adjoint all run:

* ``squeezed``: the 27-node cube squeezed 15% along x with noise, bounds
  0.93 / 1.07 (every element clamped somewhere);
* ``inverted``: the same cube at the production bounds 0.05 / 20 with one
  element pushed through itself (det F < 0 at a Tet4 or at Hex8 Gauss points);
* ``compressed`` / ``stretched``: the per-axis stretches (0.03, 0.8, 1.2) and
  (21, 1.1, 0.9) of the reference's tests/test_activation_gradients.py:183-209
  at bounds 0.05 / 20.

Compared: element forces with Rayleigh damping (1e-5 relative to the largest
force), assembled values (1e-5) and block-Jacobi inverses (3e-5), a converged
step with both CG kernels (increment 2e-4, positions 1e-6), and the gradients
of a differentiable step w.r.t. positions, velocities, external forces and
activations against the oracle's reverse products (2e-3).
"""

import numpy as np
import pytest
import torch

import oracle
from helpers import cube

pytestmark = pytest.mark.gpu

DEV = "cuda:0"
DT = 0.002
MATERIAL = (500.0, 1000.0, 0.3)
BOUNDS = {"squeezed": (0.93, 1.07), "inverted": (0.05, 20.0), "compressed": (0.05, 20.0), "stretched": (0.05, 20.0)}
CASES = list(BOUNDS)


def rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30))


def np_(t):
    return t.detach().cpu().numpy().astype(np.float64)


def deformed(model, kind, case, rng):
    """Positions (float32-representable) and velocities of a clamp case."""
    q = model.host["particle_q"].astype(np.float64).copy()
    lo = q.min(axis=0)
    if case == "squeezed":
        c = q.mean(axis=0)
        q = c + (q - c) * np.array([0.85, 1.075, 1.0]) + rng.normal(0.0, 0.004, q.shape)
    elif case == "compressed":
        q = lo + (q - lo) * np.array([0.03, 0.8, 1.2])
    elif case == "stretched":
        q = lo + (q - lo) * np.array([21.0, 1.1, 0.9])
    else:
        q = q + rng.normal(0.0, 0.002, q.shape)
        if kind == "tet":
            # Node 1 belongs to two tets; reflect it through the plane of
            # tet 1's other corners so that tet 1 is inverted.
            ids = model.host["tet_indices"][1]
            n = int(ids[0]) if ids[0] == 1 else 1
            others = [int(k) for k in ids if k != n]
            a, b, c = q[others]
            nrm = np.cross(b - a, c - a)
            nrm /= np.linalg.norm(nrm)
            q[n] -= 2.0 * np.dot(q[n] - a, nrm) * nrm
        else:
            # Corner node 0 of hex 0 pulled past the element's centre.
            ids = model.host["hex_indices"][0]
            centre = q[ids].mean(axis=0)
            q[ids[0]] = q[ids[0]] + 1.7 * (centre - q[ids[0]])
    q = q.astype(np.float32).astype(np.float64)
    v = np.tile((0.03, -0.02, 0.04), (len(q), 1)) + rng.normal(0.0, 0.05, q.shape)
    return q, v.astype(np.float32).astype(np.float64)


def jacobians(model, kind, q):
    """det F of every Tet4 / Hex8 Gauss point (unclamped), FP64."""
    h = model.host
    if kind == "tet":
        out = []
        for ids, pose in zip(h["tet_indices"], h["tet_poses"].astype(np.float64)):
            grad = np.vstack((-pose.sum(axis=0), pose))
            out.append(np.linalg.det(q[ids].T @ grad))
        return np.array(out)
    from newton_solver_fem_amd.model import HEX_SIGNS

    gp = HEX_SIGNS / np.sqrt(3.0)
    s = HEX_SIGNS
    out = []
    for ids, invs in zip(h["hex_indices"], h["hex_inv_J0"].astype(np.float64)):
        for g in range(8):
            xi, eta, zeta = gp[g]
            dxi = np.stack([s[:, 0] * (1 + s[:, 1] * eta) * (1 + s[:, 2] * zeta),
                            (1 + s[:, 0] * xi) * s[:, 1] * (1 + s[:, 2] * zeta),
                            (1 + s[:, 0] * xi) * (1 + s[:, 1] * eta) * s[:, 2]], axis=1) * 0.125
            out.append(np.linalg.det(q[ids].T @ (dxi @ invs[g])))
    return np.array(out)


def setup(kind, case, differentiable=False, **extra):
    from newton_solver_fem_amd.solver import SolverFEM

    lo, hi = BOUNDS[case]
    model = cube(kind, material=MATERIAL, gravity=(0.0, 0.0, -9.81), device=DEV)
    kw = dict(dt=DT, min_stretch=lo, max_stretch=hi, enable_tet_activations=(kind == "tet"))
    kw.update(extra)
    s = SolverFEM(model, solver_type="cg", preconditioner_type="block_diag", self_contact_ke=0.0,
                  differentiable=differentiable, **kw)
    names = dict(dt="dt", min_stretch="min_stretch", max_stretch="max_stretch",
                 enable_tet_activations="enable_tet_activations", linear_solver_tolerance="tolerance",
                 linear_solver_maxiter="maxiter")
    o = oracle.OracleSolver(model.host, bits=64, gravity=model.gravity,
                            **{names[k]: v for k, v in kw.items() if k in names})
    return model, s, o


def engaged(model, kind, q, v, o):
    """The clamp changes the oracle's force somewhere (so the case binds)."""
    term = oracle.TERM_TET if kind == "tet" else oracle.TERM_HEX
    free = oracle.OracleSolver(model.host, bits=64, gravity=model.gravity, min_stretch=-1.0)
    a, b = o.eval_terms(term, q, v), free.eval_terms(term, q, v)
    return rel(a, b) > 1e-3


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("kind", ["tet", "hex"])
def test_clamped_forces_and_assembly_match_oracle(kind, case):
    print(f"case clamp forces/assembly {kind} {case}", flush=True)
    rng = np.random.default_rng(31)
    model, s, o = setup(kind, case)
    q, v = deformed(model, kind, case, rng)
    dets = jacobians(model, kind, q)
    if case == "inverted":
        assert dets.min() < 0.0, dets.min()
    assert engaged(model, kind, q, v, o)
    st = model.state()
    st.particle_q.copy_(torch.tensor(q, dtype=torch.float32))
    st.particle_qd.copy_(torch.tensor(v, dtype=torch.float32))
    term = oracle.TERM_TET if kind == "tet" else oracle.TERM_HEX
    gpu_f = np_(s.eval_tetrahedral_forces(model, None, st) if kind == "tet" else
                s.eval_hexahedral_forces(model, None, st))
    ref_f = o.eval_terms(term, q, v)
    err_f = rel(gpu_f, ref_f)
    s.refresh_tangent(st.particle_q)
    o.refresh_tangent(q)
    _, _, vals, minv = o.matrix()
    err_a = rel(np_(s.A_bsr.values), vals)
    err_m = rel(np_(s._pc_inv_diag), minv)
    print(f"  det F min {dets.min():.3e}: force {err_f:.2e}, values {err_a:.2e}, inverse {err_m:.2e}", flush=True)
    assert err_f < 1e-5
    assert err_a < 1e-5
    assert err_m < 3e-5


@pytest.mark.parametrize("variant", ["resident", "streamed"])
@pytest.mark.parametrize("case", ["squeezed", "inverted"])
@pytest.mark.parametrize("kind", ["tet", "hex"])
def test_clamped_step_matches_oracle(kind, case, variant):
    """Three converged steps (tol 1e-7) from a clamped state, both CG kernels."""
    print(f"case clamp step {kind} {case} {variant}", flush=True)
    rng = np.random.default_rng(37)
    model, s, o = setup(kind, case, linear_solver_tolerance=1e-7, linear_solver_maxiter=300)
    s.set_cg_variant(variant)
    assert s.launch_config()["cg_variant"] == variant
    q, v = deformed(model, kind, case, rng)
    st0, st1 = model.state(), model.state()
    st0.particle_q.copy_(torch.tensor(q, dtype=torch.float32))
    st0.particle_qd.copy_(torch.tensor(v, dtype=torch.float32))
    act = rng.uniform(0.0, 0.3, model.tet_count) if kind == "tet" else None
    ctrl = model.control()
    if act is not None:
        ctrl.tet_activations = torch.tensor(act, dtype=torch.float32, device=DEV)
    for k in range(3):
        s.refresh_tangent(st0.particle_q)
        o.refresh_tangent(q)
        s.step(st0, st1, ctrl, None, DT)
        out = o.step(q, v, activations=act)
        e_f, e_dv = rel(np_(st0.particle_f), out["f"]), rel(np_(s.dv), o.dv)
        e_q, e_v = rel(np_(st1.particle_q), out["q"]), rel(np_(st1.particle_qd), out["qd"])
        print(f"  step {k}: rhs {e_f:.2e}, dv {e_dv:.2e}, q {e_q:.2e}, qd {e_v:.2e}", flush=True)
        assert e_f < 1e-5 and e_dv < 2e-4 and e_v < 2e-4 and e_q < 1e-6, k
        q, v = np_(st1.particle_q), np_(st1.particle_qd)
        st0, st1 = st1, st0


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("kind", ["tet", "hex"])
def test_clamped_step_gradients_match_oracle(kind, case):
    """One differentiable step (checked solves) with the clamp binding: the
    forward state and every input gradient vs the oracle's reverse products
    (clamp adjoint kernels.py:447-481)."""
    print(f"case clamp diff {kind} {case}", flush=True)
    rng = np.random.default_rng(41)
    model, s, o = setup(kind, case, differentiable=True, linear_solver_maxiter=256)
    q0, v0 = deformed(model, kind, case, rng)
    s.refresh_tangent(torch.tensor(q0, dtype=torch.float32, device=DEV))
    o.refresh_tangent(q0)
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
    st_in, st_out = model.state(), model.state()
    st_in.particle_q, st_in.particle_qd = tq, tv
    s.step(st_in, st_out, ctrl, None, DT, external_forces=te)
    e_q, e_v = rel(np_(st_out.particle_q), qo), rel(np_(st_out.particle_qd), vo)
    loss = (st_out.particle_q * torch.tensor(gq_out, dtype=torch.float32, device=DEV)).sum() + (
        st_out.particle_qd * torch.tensor(gv_out, dtype=torch.float32, device=DEV)).sum()
    loss.backward()
    res = s.check_differentiation()
    u = gv_out + DT * gq_out
    errs = {"q": rel(np_(tq.grad) - gq_out, ref["q"] - gq_out), "qd": rel(np_(tv.grad) - u, ref["qd"] - u),
            "ext": rel(np_(te.grad), ref["ext"])}
    if act is not None:
        errs["act"] = rel(np_(ctrl.tet_activations.grad), ref["act"])
    print(f"  forward q {e_q:.2e} qd {e_v:.2e}; gradients " + ", ".join(f"{k} {e:.2e}" for k, e in errs.items()),
          flush=True)
    assert e_q < 1e-6 and e_v < 1e-4
    assert res["forward_relative_residual"] < 2e-5 and res["backward_relative_residual"] < 2e-5
    for name, e in errs.items():
        assert e < 2e-3, (name, e)
