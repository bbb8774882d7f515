"""Prepared affine mechanics of one non-differentiable substep.

Mirrors the reference's ``PreparedFEMStep`` (``prepared_step.py:64-302``,
obtained through ``SolverFEM.prepare_step``, ``solver.py:483-500``): the force
terms and the elastic predictor are evaluated once at ``prepare``; the
right-hand side for additional loads, the constrained operator and
preconditioner actions, the solve and the state reconstruction are then
resident device operations (HIP kernels behind ``libfem_hip.so``) that can be
recorded into a HIP graph.
"""

from __future__ import annotations

import math

import torch

from . import _lib


class PreparedFEMStep:
    """A prepared equation ``operator * delta_v = rhs(external_forces)``.

    ``base_rhs`` is the zero-additional-load impulse [N s] including the
    elastic predictor; ``fixed_mask`` and ``diagonal_blocks`` describe the
    constrained operator.  Input positions and velocities are copied into
    persistent storage, so trial reconstruction cannot overwrite the accepted
    state.  One workspace is reused per solver; preparing again invalidates the
    previous preparation.  Refresh the tangent separately when changing the
    linearisation or the timestep, as for ``SolverFEM.step``.
    """

    def __init__(self, solver):
        self.solver, self.model = solver, solver.model
        self.device = solver.device
        self._state = self.model.state()
        P = self.model.particle_count
        z = lambda: torch.zeros((P, 3), dtype=torch.float32, device=self.device)
        self.base_rhs = z()
        self._rhs = z()
        self.velocity_increment = z()
        self.fixed_mask = torch.zeros(P, dtype=torch.int32, device=self.device)
        if solver.fixed_particle_indices.size:
            self.fixed_mask[torch.as_tensor(solver.fixed_particle_indices, dtype=torch.long,
                                            device=self.device)] = 1
        self.diagonal_blocks = torch.zeros((P, 3, 3), dtype=torch.float32, device=self.device)
        self.unconstrained_matrix = solver.A_bsr
        self._revision = None
        self.dt = None

    # ------------------------------------------------------------------
    @staticmethod
    def check_supported(solver):
        if solver.differentiable:
            raise RuntimeError("Prepared FEM steps are non-differentiable and cannot run inside a tape")
        if solver._streaming or solver._component_scaling is not None:
            raise ValueError("Prepared FEM steps require assembled, unscaled particle coordinates")

    def _check_current(self):
        self.check_supported(self.solver)
        if self._revision is None:
            raise RuntimeError("Prepare FEM mechanics before using this workspace")
        if self._revision != self.solver._revision or self.dt != self.solver.dt:
            raise RuntimeError("FEM mechanics changed; prepare the step again")

    def _vectors(self, t, name):
        if (not isinstance(t, torch.Tensor) or t.dtype != torch.float32 or tuple(t.shape) != (self.model.particle_count, 3)
                or t.device != self.device or not t.is_contiguous()):
            raise ValueError(f"{name} must be a particle-sized (P, 3) float32 tensor on the model device")

    def _call(self, fn, *args):
        _lib.check(getattr(self.solver._lib, fn)(self.solver._sys, *args), fn)

    @staticmethod
    def _stream(device):
        return torch.cuda.current_stream(device).cuda_stream

    # ------------------------------------------------------------------
    def prepare(self, state_in, control, contacts, dt, external_forces=None):
        self.check_supported(self.solver)
        if torch.cuda.is_current_stream_capturing():
            raise RuntimeError("Prepare FEM mechanics before CUDA graph capture")
        dt = float(dt)
        if not math.isfinite(dt) or dt <= 0 or dt != self.solver.dt:
            raise ValueError("Prepared timestep must be finite, positive and match the assembled tangent timestep")
        self._vectors(state_in.particle_q, "particle_q")
        self._vectors(state_in.particle_qd, "particle_qd")
        if external_forces is not None:
            self._vectors(external_forces, "external_forces")
        solver = self.solver
        solver._sync_params()
        self._revision = None
        self._state.particle_q.copy_(state_in.particle_q)
        self._state.particle_qd.copy_(state_in.particle_qd)
        act = None
        if solver.enable_tet_activations:
            act = (control if control is not None else self.model.control()).tet_activations.contiguous()
        st = self._stream(self.device)
        # forces, h * load and predictor at the owned copy (fixed particles restored in it)
        self._call("fem_step_rhs", self._state.particle_q.data_ptr(), self._state.particle_qd.data_ptr(),
                   self.base_rhs.data_ptr(), None if external_forces is None else external_forces.data_ptr(),
                   None if act is None else act.data_ptr(), dt, st)
        self._call("fem_diagonal_blocks", self.diagonal_blocks.data_ptr(), st)
        self.velocity_increment.copy_(solver.velocity_increment)
        self.dt, self._revision = dt, solver._revision
        return self

    def rhs(self, external_forces=None, out=None):
        """The constrained impulse with additional nodal loads [N] (fixed rows zero)."""
        self._check_current()
        out = self._rhs if out is None else out
        self._vectors(out, "out")
        if external_forces is not None:
            self._vectors(external_forces, "external_forces")
            if external_forces.data_ptr() == out.data_ptr():
                raise ValueError("RHS output must not alias its external-force input")
        self._call("fem_add_load", self.base_rhs.data_ptr(),
                   None if external_forces is None else external_forces.data_ptr(), self.dt, out.data_ptr(),
                   self._stream(self.device))
        return out

    def apply_operator(self, vector, out):
        """Constrained momentum action (A, or PAP + I - P with fixed particles)."""
        self._check_current()
        self._vectors(vector, "vector")
        self._vectors(out, "out")
        if vector.data_ptr() == out.data_ptr():
            raise ValueError("Operator action requires distinct input and output storage")
        self._call("fem_apply_operator", vector.data_ptr(), out.data_ptr(), self._stream(self.device))
        return out

    def apply_preconditioner(self, vector, out):
        """Block-Jacobi preconditioner action."""
        self._check_current()
        self._vectors(vector, "vector")
        self._vectors(out, "out")
        if vector.data_ptr() == out.data_ptr():
            raise ValueError("Preconditioner action requires distinct input and output storage")
        self._call("fem_apply_preconditioner", vector.data_ptr(), out.data_ptr(), self._stream(self.device))
        return out

    def solve(self, rhs, out=None, *, tolerance=None):
        """The step's solve on ``rhs``; changes only this workspace's numerical guess."""
        self._check_current()
        self._vectors(rhs, "rhs")
        solver = self.solver
        saved = solver._params.tolerance
        if tolerance is not None:
            if not (math.isfinite(tolerance) and tolerance > 0):
                raise ValueError("linear solver tolerance must be finite and positive")
            solver._params.tolerance = float(tolerance)
            solver._lib.fem_set_params(solver._sys, solver._params_ref())
        try:
            self._call("fem_solve_increment", rhs.data_ptr(), self.velocity_increment.data_ptr(),
                       self._stream(self.device))
        finally:
            if tolerance is not None:
                solver._params.tolerance = saved
                solver._lib.fem_set_params(solver._sys, solver._params_ref())
        if out is not None:
            self._vectors(out, "out")
            out.copy_(self.velocity_increment)
            return out
        return self.velocity_increment

    def reconstruct(self, delta_v, state_out):
        """Particle q/qd from the prepared state and ``delta_v``, fixed supports enforced."""
        self._check_current()
        self._vectors(delta_v, "delta_v")
        self._vectors(state_out.particle_q, "particle_q")
        self._vectors(state_out.particle_qd, "particle_qd")
        if state_out.particle_q.data_ptr() == self._state.particle_q.data_ptr():
            raise ValueError("Reconstruction must not overwrite the prepared input")
        self._call("fem_reconstruct", delta_v.data_ptr(), self._state.particle_q.data_ptr(),
                   self._state.particle_qd.data_ptr(), state_out.particle_q.data_ptr(),
                   state_out.particle_qd.data_ptr(), self.dt, self._stream(self.device))
        return state_out
