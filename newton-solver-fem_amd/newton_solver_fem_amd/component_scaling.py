"""Equation scaling of disconnected components (reference component_scaling.py).

Every equation of a mechanically disconnected component is divided by the
component's total mass: the solve uses ``S A``, ``M^-1 S^-1`` and ``S b`` with
``S = diag(1 / m_component)``.  ``S`` is uniform inside a component, so the
scaled matrix stays symmetric and its exact solution is the physical velocity
increment, while a finite CG budget now weighs a tiny separated piece like a
large one.  The HIP library forms the factors (``fem_refresh_component_scaling``)
and writes the scaled copies before every solve (``k_scale_system``); the
storage here is caller-owned and persistent, so captured graphs stay valid.
"""

from __future__ import annotations

import torch

from . import _lib


class ComponentScaling:
    """Scaled copies of the solver's matrix, inverse diagonal and RHS
    (reference component_scaling.py:52-122)."""

    def __init__(self, solver, labels: torch.Tensor, capacity: int):
        self.solver = solver
        self.model = solver.model
        self.labels = labels
        self.capacity = int(capacity)
        dev = solver.device
        P = self.model.particle_count
        block_capacity = solver._A_storage[2].shape[0]
        self.total_mass = torch.zeros(self.capacity, dtype=torch.float64, device=dev)
        self.scale = torch.ones(max(P, 1), dtype=torch.float32, device=dev)
        self.rhs = torch.zeros(max(P, 1), 3, dtype=torch.float32, device=dev)
        self.inverse = torch.zeros(max(P, 1), 3, 3, dtype=torch.float32, device=dev)
        self._values = torch.zeros(block_capacity, 3, 3, dtype=torch.float32, device=dev)
        _lib.check(
            solver._lib.fem_bind_component_scaling(
                solver._sys, self.capacity, self.total_mass.data_ptr(), self.scale.data_ptr(),
                self._values.data_ptr(), self.inverse.data_ptr(), self.rhs.data_ptr()),
            "prepare_component_scaling",
        )
        self.refresh()

    @property
    def matrix(self):
        """Scaled matrix: the solver's offsets and columns, own values."""
        from .solver import BSRMatrix

        A = self.solver.A_bsr
        return BSRMatrix(A.offsets, A.columns, self._values[: A.nnz], A.nrow, A.ncol)

    def refresh(self):
        """Per-component total masses and row factors from the current labels
        and model masses (reference component_scaling.py:95-109)."""
        from .solver import _stream

        _lib.check(
            self.solver._lib.fem_refresh_component_scaling(
                self.solver._sys, self.labels.data_ptr(), self.model.particle_mass.data_ptr(),
                _stream(self.solver.device)),
            "refresh_component_scaling",
        )
