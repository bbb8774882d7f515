"""Differentiable substep on the HIP library (reference ``differentiation.py``).

The reference differentiates one linearised step with Warp's tape
(differentiation.py:289-394): the matrix is a fixed input (the current
assembly, ``MatrixSnapshot`` :149-242), the forward solve and the adjoint
solve are checked block-Jacobi CG solves to 1e-8 from a zero start, and the
tape carries the adjoint through the force kernels, ``form_rhs`` (the
elastic predictor) and ``update_state``.  Here the same linearisation is a
``torch.autograd.Function`` over the HIP C-ABI:

* forward: ``fem_step_rhs`` (h * sum of terms + h * ext - sum_j A_ij (v_j - v_i))
  then ``fem_solve`` (kind 0) -> dv; ``v' = v + dv`` and ``q' = q + h v'`` are
  plain torch operations;
* backward: ``fem_solve`` (kind 1) gives z = A^-1 dv_bar (A is symmetric), then
  ``fem_vjp`` returns the input gradients (positions, velocities,
  activations, external forces = h z, and -- beyond the reference, which
  keeps materials fixed -- Tet4 (mu, lambda, k_damp) including the matrix
  terms -z^T (dA/dmu)(dv + v)).

``rollout`` chains steps and, with ``checkpoint_every=k``, keeps only every
k-th state for the backward pass and recomputes the segments in between
(the frame checkpoints of the reference's ``BearProblem.forward/backward``,
examples/diffsim/_bear_training.py:254-310).
"""

from __future__ import annotations

import ctypes

import torch

from . import _lib
from .model import State

CHECKED_TOLERANCE = 1.0e-8  # MatrixSnapshot's forward/backward cg tolerance (differentiation.py:162-177)


def _stream(solver):
    return torch.cuda.current_stream(solver.device).cuda_stream


def _ptr(t):
    return None if t is None else t.data_ptr()


def _fixed_mask(solver):
    m = getattr(solver, "_fixed_mask_t", None)
    if m is None:
        m = torch.zeros(solver.model.particle_count, 1, dtype=torch.bool, device=solver.device)
        if solver.fixed_particle_indices.size:
            m[torch.as_tensor(solver.fixed_particle_indices, dtype=torch.long, device=solver.device)] = True
        solver._fixed_mask_t = m
    return m


def prepare(solver):
    """Validate the model for differentiation and pin the current matrix revision.

    Reference ``prepare`` (differentiation.py:244-266): no streaming topology,
    no particle or surface contacts (only the analytic ground), fixed model
    arrays.  The fixed matrix of a recorded step is the current assembly.
    Fixed particles (allowed only with allow_fixed_in_differentiation) use
    the projected operator in both solves.
    """
    if solver.model.particle_ke > 0.0:
        raise NotImplementedError("Differentiation supports analytic ground contact; disable particle/surface contacts")
    if torch.cuda.is_current_stream_capturing():
        raise RuntimeError("Call prepare_differentiation() before CUDA graph capture")
    if getattr(solver, "_q_asm", None) is None:
        solver._q_asm = solver.model.particle_q.detach().clone()
    solver._diff_revision = solver._revision
    return solver._diff_revision


def check(solver, *, reset: bool = True) -> dict:
    """Read (and by default clear) the checked-solve residual maxima; raise on failure.

    Reference ``check`` (differentiation.py:269-279): relative residuals of the
    forward and adjoint solves, and failure when any solve's residual exceeds
    its budget 2e-5 ||b|| + 1e-10 or is not finite.
    """
    if torch.cuda.is_current_stream_capturing():
        raise RuntimeError("Check differentiable residuals after launching the CUDA graph")
    out = (ctypes.c_double * 4)()
    _lib.check(solver._lib.fem_diff_stats(solver._sys, _stream(solver), out, 1 if reset else 0), "fem_diff_stats")
    result = {"forward_relative_residual": float(out[0]), "backward_relative_residual": float(out[1])}
    solver.last_differentiation_residuals = result
    vals = [out[k] for k in range(4)]
    if not all(v == v and abs(v) != float("inf") for v in vals) or vals[2] > 1.0 or vals[3] > 1.0:
        raise RuntimeError(f"FEM differentiable linear solve did not converge: {result}")
    return result


def _solve(solver, b, kind):
    x = torch.empty_like(b)
    _lib.check(solver._lib.fem_solve(solver._sys, b.data_ptr(), x.data_ptr(), CHECKED_TOLERANCE, solver._maxiter,
                                     kind, _stream(solver)), "fem_solve")
    return x


def _forward_dv(solver, dt, q, v, act, ext):
    rhs = torch.empty_like(q)
    _lib.check(solver._lib.fem_step_rhs(solver._sys, q.data_ptr(), v.data_ptr(), rhs.data_ptr(), _ptr(ext), _ptr(act),
                                        float(dt), _stream(solver)), "fem_step_rhs")
    return _solve(solver, rhs, 0)


class _VelocityIncrement(torch.autograd.Function):
    """dv = A^-1 rhs(q, v, act, ext; materials) with A the fixed current assembly."""

    @staticmethod
    def forward(ctx, solver, dt, q, v, act, ext, mat):
        q, v = q.detach().contiguous(), v.detach().contiguous()
        act = None if act is None else act.detach().contiguous()
        ext = None if ext is None else ext.detach().contiguous()
        dv = _forward_dv(solver, dt, q, v, act, ext)
        ctx.solver, ctx.dt, ctx.revision = solver, float(dt), solver._revision
        ctx.has = (act is not None, ext is not None, mat is not None)
        ctx.save_for_backward(q, v, act if act is not None else q.new_empty(0), dv)
        return dv

    @staticmethod
    def backward(ctx, gdv):
        solver, dt = ctx.solver, ctx.dt
        if solver._revision != ctx.revision:
            raise RuntimeError("FEM model/matrix changed after recording; rerun the forward before backward")
        q, v, act, dv = ctx.saved_tensors
        has_act, has_ext, has_mat = ctx.has
        act = act if has_act else None
        g = gdv.detach()
        if solver.fixed_particle_indices.size:  # dv is zero on fixed rows: project the adjoint
            g = torch.where(_fixed_mask(solver), torch.zeros_like(g), g)
        z = _solve(solver, g.contiguous(), 1)
        qb, vb = torch.empty_like(q), torch.empty_like(v)
        want_act = has_act and ctx.needs_input_grad[4]
        want_mat = has_mat and ctx.needs_input_grad[6]
        ab = torch.empty(solver.model.tet_count, dtype=torch.float32, device=q.device) if want_act else None
        mb = torch.empty(solver.model.tet_count, 3, dtype=torch.float32, device=q.device) if want_mat else None
        _lib.check(
            solver._lib.fem_vjp(solver._sys, q.data_ptr(), v.data_ptr(), _ptr(act), z.data_ptr(), dv.data_ptr(),
                                solver._q_asm.data_ptr(), float(dt), qb.data_ptr(), vb.data_ptr(), _ptr(ab), _ptr(mb),
                                _stream(solver)),
            "fem_vjp",
        )
        eb = z * dt if (has_ext and ctx.needs_input_grad[5]) else None
        if not solver.defer_differentiation_checks:
            check(solver)
        return None, None, qb, vb, ab, eb, mb


def differentiable_step(solver, state_in: State, state_out: State, control, contacts, dt: float,
                        external_forces: torch.Tensor | None = None) -> State:
    """Reference ``step`` (differentiation.py:282-394) for ``SolverFEM(differentiable=True)``.

    When any input requires a gradient (and grad mode is on), the outputs are
    new tensors attached to the autograd graph and ``state_out.particle_q`` /
    ``particle_qd`` are rebound to them; otherwise ``state_out`` is written in
    place.
    """
    if dt != solver.dt:
        raise ValueError("Differentiable dt must match the fixed matrix assembly timestep")
    if contacts is not None:
        raise NotImplementedError("Differentiation supports only the analytic ground contact law")
    prepare(solver)
    q, v = state_in.particle_q, state_in.particle_qd
    fixed = solver.fixed_particle_indices.size > 0
    if fixed:  # FixedParticles.restore(state_in), differentiably
        mask = _fixed_mask(solver)
        q = torch.where(mask, solver._anchor, q)
        v = torch.where(mask, torch.zeros_like(v), v)
    act = None
    if solver.enable_tet_activations:
        act = (control if control is not None else solver.model.control()).tet_activations
    mat = solver.tet_materials if solver.tet_materials.requires_grad else None
    taped = torch.is_grad_enabled() and any(t is not None and t.requires_grad for t in (q, v, act, external_forces, mat))
    if taped:
        if state_in is state_out or state_in.particle_q is state_out.particle_q:
            raise ValueError("Use unique input/output states for every taped step")
        dv = _VelocityIncrement.apply(solver, dt, q, v, act, external_forces, mat)
        v_out = v + dv
        q_out = q + dt * v_out
        if fixed:
            q_out = torch.where(mask, solver._anchor, q_out)
            v_out = torch.where(mask, torch.zeros_like(v_out), v_out)
        state_out.particle_qd = v_out
        state_out.particle_q = q_out
    else:
        with torch.no_grad():
            dv = _forward_dv(solver, dt, q.contiguous(), v.contiguous(),
                             None if act is None else act.contiguous(),
                             None if external_forces is None else external_forces.contiguous())
            torch.add(v, dv, out=state_out.particle_qd)
            torch.add(q, state_out.particle_qd, alpha=dt, out=state_out.particle_q)
            if fixed:
                state_out.particle_q.copy_(torch.where(mask, solver._anchor, state_out.particle_q))
                state_out.particle_qd.masked_fill_(mask, 0.0)
    with torch.no_grad():
        solver.dv.copy_(dv)
    if not solver.defer_differentiation_checks:
        check(solver)
    return state_out


def _per_step(x, k, per_step_dim):
    if x is None:
        return None
    if callable(x):
        return x(k)
    return x[k] if x.dim() == per_step_dim else x


def rollout(solver, q: torch.Tensor, qd: torch.Tensor, steps: int, *, activations=None, external_forces=None,
            checkpoint_every: int = 0, after_step=None):
    """Differentiable rollout of ``steps`` substeps from (q, qd); returns the final (q, qd).

    ``activations``: None, a (T,) tensor for every step, a (steps, T) tensor
    or a callable ``k -> (T,)``; ``external_forces`` likewise with (P, 3) /
    (steps, P, 3).  ``checkpoint_every=k > 0`` stores only every k-th state
    for the backward pass and recomputes each k-step segment from it
    (torch.utils.checkpoint), trading one extra forward for O(steps / k)
    memory instead of O(steps).  The last segment is recorded directly: the
    backward pass would recompute it first, so keeping its states raises no
    memory peak and saves that segment's recomputation.  ``after_step(q, qd) -> (q, qd)`` is applied
    after every substep (differentiably), e.g. the support projection the
    reference uses for zero-mass pinned nodes (tests/diffsim_optimize.py:29-45).
    """
    dt = solver.dt
    model = solver.model

    def run(q0, v0, k0, k1, *tensors):
        a_all, e_all = tensors if tensors else (None, None)
        st = State(q0, v0, None)
        for k in range(k0, k1):
            out = State(torch.empty_like(q0), torch.empty_like(q0), None)
            ctrl = None
            a = _per_step(activations if a_all is None else a_all, k, 2)
            if a is not None:
                ctrl = model.control()
                ctrl.tet_activations = a
            e = _per_step(external_forces if e_all is None else e_all, k, 3)
            differentiable_step(solver, st, out, ctrl, None, dt, e)
            qn, vn = out.particle_q, out.particle_qd
            if after_step is not None:
                qn, vn = after_step(qn, vn)
            st = State(qn, vn, None)
        return st.particle_q, st.particle_qd

    if checkpoint_every <= 0 or not torch.is_grad_enabled():
        return run(q, qd, 0, steps)
    import torch.utils.checkpoint as ckpt

    extra = []
    for x in (activations, external_forces):
        extra.append(x if isinstance(x, torch.Tensor) else None)
    for k0 in range(0, steps, checkpoint_every):
        k1 = min(steps, k0 + checkpoint_every)
        if k1 == steps:
            q, qd = run(q, qd, k0, k1, *extra)
        else:
            q, qd = ckpt.checkpoint(run, q, qd, k0, k1, *extra, use_reentrant=False)
    return q, qd
