"""``SolverFEM``: the linearly implicit Tet4/Hex8 substep on MI355X.

Host-side mirror of the reference's ``SolverFEM``
(src/newton_solver_fem/_src/solvers/fem/solver.py:132-1701) for its
substep path: same constructor arguments, ``step`` / ``refresh_tangent``
signatures, side effects and error behaviour.  All numerics run in
``libfem_hip.so`` (hand-written HIP for gfx950) through the C-ABI in
``include/fem_hip.h``; this module only validates arguments, owns the
persistent device buffers and forwards raw pointers plus the current
torch stream.  There is no CPU or torch fallback.
"""

from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from .model import Control, Model, State

# solver_type / preconditioner_type -> fem_set_solver codes (solver.py:1142-1173, :1219-1252).
SOLVER_TYPES = {"cg": 0, "bicgstab": 1, "cr": 2, "gmres": 3}
PRECONDITIONER_TYPES = {"block_diag": 0, "diag": 1, "diag_abs": 2, "id": 3}
GMRES_RESTART = 31  # warp.optim.linear.gmres default


class BlockJacobiOperator:
    """The block-Jacobi preconditioner ``M_bsr`` of the reference
    (solver.py:1175-1201: a warp LinearOperator over ``_pc_inv_diag``).

    ``matvec(x, y, z, alpha, beta)`` writes ``z = alpha M^-1 x + beta y``
    (block_jacobi_mv_kernel kernels.py:1852-1868) on the HIP device.
    """

    def __init__(self, solver):
        self._solver = solver
        n = solver.model.particle_count
        self.shape = (3 * n, 3 * n)
        self.dtype = torch.float32
        self.device = solver.device

    def matvec(self, x, y, z, alpha, beta):
        s = self._solver
        n = s.model.particle_count
        for name, t in (("x", x), ("y", y), ("z", z)):
            _check_vec3(name, t, n, s.device)
        _lib.check(s._lib.fem_block_jacobi_mv(s._sys, x.data_ptr(), y.data_ptr(), z.data_ptr(), float(alpha),
                                              float(beta), _stream(s.device)), "M_bsr.matvec")
        return z


@dataclass
class BSRMatrix:
    """Block-sparse (3x3) matrix view over device tensors (warp BsrMatrix subset)."""

    offsets: torch.Tensor  # [nrow+1] int32
    columns: torch.Tensor  # [nnz] int32
    values: torch.Tensor  # [nnz, 3, 3] float32
    nrow: int
    ncol: int

    @property
    def nnz(self) -> int:
        return int(self.columns.numel())

    def dense(self) -> np.ndarray:
        """FP64 dense copy (small matrices; test helper)."""
        off = self.offsets.cpu().numpy()
        cols = self.columns.cpu().numpy()
        vals = self.values.detach().cpu().numpy().astype(np.float64)
        out = np.zeros((3 * self.nrow, 3 * self.ncol))
        for r in range(self.nrow):
            for k in range(off[r], off[r + 1]):
                c = cols[k]
                out[3 * r : 3 * r + 3, 3 * c : 3 * c + 3] += vals[k]
        return out


def _stream(device: torch.device) -> int:
    return torch.cuda.current_stream(device).cuda_stream


def _ptr(t: torch.Tensor | None) -> int | None:
    return None if t is None else t.data_ptr()


def _check_vec3(name: str, t, count: int, device: torch.device):
    if (
        not isinstance(t, torch.Tensor)
        or t.dtype != torch.float32
        or tuple(t.shape) != (count, 3)
        or t.device != device
        or not t.is_contiguous()
    ):
        raise ValueError(f"{name} must be a contiguous float32 tensor of shape ({count}, 3) on {device}")


class SolverFEM:
    """Linearly implicit backward-Euler FEM solver (Tet4 + Hex8, SNH).

    Constructor arguments follow the reference (solver.py:212-240).  The
    linear solve runs ``solver_type`` ``"cg"`` (the persistent PCG),
    ``"bicgstab"``, ``"cr"`` or ``"gmres"`` with ``preconditioner_type``
    ``"block_diag"``, ``"diag"``, ``"diag_abs"`` or ``"id"``.  A ``contacts``
    argument with soft contacts (:class:`Contacts`, Newton's contact list)
    adds the particle-shape soft-contact term to the step.
    """

    def __init__(
        self,
        model: Model,
        dt: float = 1.0 / 60.0,
        mass: float = 1.0,
        preconditioner_type: str = "diag",
        solver_type: str = "bicgstab",
        linear_solver_maxiter: int = 50,
        min_stretch: float = 0.05,
        max_stretch: float = 20.0,
        ground_plane=None,
        ground_ke: float = 1.0e5,
        ground_kd: float = 1.0e2,
        ground_kf: float = 1.0e3,
        ground_mu: float = 0.5,
        torque_stiffness: float = 0.0,
        torque_damping: float = 0.0,
        spring_rest_direction=None,
        self_contact_ke: float = 1.0e3,
        self_contact_kd: float = 1.0e2,
        linear_damping: float = 0.0,
        differentiable: bool = False,
        enable_tet_activations: bool = False,
        defer_differentiation_checks: bool = False,
        tet_volume_weights=None,
        fixed_particle_indices=None,
        linear_solver_tolerance: float = 1.0e-2,
        linear_solver_use_cuda_graph: bool = True,
        instance_offsets=None,
        allow_fixed_in_differentiation: bool = False,
    ):
        # Argument validation, reference solver.py:241-261.
        if not math.isfinite(linear_solver_tolerance) or linear_solver_tolerance <= 0:
            raise ValueError("linear_solver_tolerance must be finite and positive")
        fixed = np.asarray([] if fixed_particle_indices is None else fixed_particle_indices)
        if fixed.ndim != 1 or (fixed.size and fixed.dtype.kind not in "iu"):
            raise ValueError("fixed_particle_indices must be a one-dimensional integer sequence")
        if fixed.size and (fixed.min() < 0 or fixed.max() >= model.particle_count):
            raise ValueError("fixed_particle_indices contains an out-of-range node")
        fixed = np.unique(fixed).astype(np.int32)
        if fixed.size and differentiable and not allow_fixed_in_differentiation:
            # Reference behaviour (solver.py:250-251).  allow_fixed_in_differentiation
            # enables the projected forward and adjoint solves instead.
            raise ValueError("fixed_particle_indices is not supported with differentiable=True")
        if not (isinstance(dt, (int, float)) and math.isfinite(dt) and dt > 0):
            raise ValueError("dt must be finite and positive")
        if solver_type not in SOLVER_TYPES:
            raise ValueError(f"Invalid solver type: {solver_type}")
        if preconditioner_type not in PRECONDITIONER_TYPES:
            raise ValueError(f"Unsupported preconditioner type {preconditioner_type}")
        if instance_offsets is not None and (solver_type, preconditioner_type) != ("cg", "block_diag"):
            raise NotImplementedError("batched instances run solver_type='cg' with preconditioner_type='block_diag'")
        # Spring bend torque (reference solver.py:390-401, eval_spring_forces :1436-1476).
        rest_dir = None
        if spring_rest_direction is not None:
            src = spring_rest_direction
            if isinstance(src, torch.Tensor):
                src = src.detach().cpu().numpy()
            rest_dir = np.ascontiguousarray(src, dtype=np.float32)
            if rest_dir.shape != (model.spring_count, 3):
                raise ValueError(f"spring_rest_direction shape {rest_dir.shape} != "
                                 f"(spring_count={model.spring_count}, 3)")
        torque_active = (float(torque_stiffness) > 0.0 and rest_dir is not None
                         and bool(np.any(np.linalg.norm(rest_dir, axis=1) >= 0.5)))
        if torque_active and differentiable:
            raise NotImplementedError("differentiable steps with spring bend torque are not supported")
        # The reference stores the contact coefficients on the model (solver.py:266-267).
        model.particle_ke = float(self_contact_ke)
        model.particle_kd = float(self_contact_kd)
        # Particle-particle contact runs only for models with a hash grid
        # (reference eval_particle_particle_forces solver.py:1477-1486); FEM
        # meshes built without one skip the term whatever self_contact_ke is.
        self._soft_key = None  # bound soft-contact arrays (none at creation)
        self._particle_contact = (model.particle_ke > 0.0 and model.particle_count > 1
                                  and getattr(model, "particle_grid", None) is not None
                                  and getattr(model, "particle_max_radius", 0.0) > 0.0)
        if self._particle_contact and differentiable:
            raise NotImplementedError("differentiable steps with particle-particle contact are not supported")
        if model.device.type != "cuda":
            raise ValueError("SolverFEM runs on the HIP device: finalize the model with device='cuda'")
        if tet_volume_weights is not None:
            w = torch.as_tensor(tet_volume_weights)
            if tuple(w.shape) != (model.tet_count,) or w.dtype != torch.float32:
                raise ValueError("tet_volume_weights must be a float32 array of shape (tet_count,)")
            if not torch.isfinite(w).all() or (w < 0).any():
                raise ValueError("tet_volume_weights must be finite and nonnegative")
            tet_volume_weights = w.to(model.device)

        self.model = model
        self.device = model.device
        self.mass = float(mass)
        self.solver_type = solver_type
        self.preconditioner_type = preconditioner_type
        self.min_stretch = float(min_stretch)
        self.max_stretch = float(max_stretch)
        self.linear_damping = float(linear_damping)
        self.differentiable = bool(differentiable)
        self.enable_tet_activations = bool(enable_tet_activations)
        self.defer_differentiation_checks = bool(defer_differentiation_checks)
        self.linear_solver_use_cuda_graph = bool(linear_solver_use_cuda_graph)
        self.linear_solver_check_every = 0
        self.last_cg_iters: int | None = None
        self.last_differentiation_residuals = None
        self.dt = float(dt)
        self.torque_stiffness = float(torque_stiffness)
        self.torque_damping = float(torque_damping)
        self.spring_rest_direction = torch.as_tensor(
            rest_dir if rest_dir is not None else np.zeros((model.spring_count, 3), np.float32), device=model.device)
        self._ground_plane = None if ground_plane is None else tuple(float(g) for g in ground_plane)
        self._ground = (float(ground_ke), float(ground_kd), float(ground_kf), float(ground_mu))
        self._tolerance = float(linear_solver_tolerance)
        self._maxiter = int(linear_solver_maxiter)
        self._tet_volume_weights = tet_volume_weights
        self.fixed_particle_indices = fixed
        self.hex_count = int(model.fem.hex_count)
        self._diff = None
        self._revision = 0
        self.topology_reallocated = False

        self._lib = _lib.load()
        host = model.host
        P = model.particle_count
        self._host_keep = {}

        def fptr(name, arr):
            a = np.ascontiguousarray(arr, dtype=np.float32)
            self._host_keep[name] = a
            return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))

        def iptr(name, arr):
            a = np.ascontiguousarray(arr, dtype=np.int32)
            self._host_keep[name] = a
            return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))

        desc = _lib.ModelDesc()
        desc.particle_count = P
        desc.particle_q = fptr("q", host["particle_q"])
        desc.particle_mass = fptr("m", host["particle_mass"])
        desc.particle_inv_mass = fptr("im", host["particle_inv_mass"])
        desc.particle_radius = fptr("r", host["particle_radius"])
        desc.particle_flags = iptr("fl", host["particle_flags"])
        desc.tet_count = model.tet_count
        desc.tet_indices = iptr("ti", host["tet_indices"])
        desc.tet_poses = fptr("tp", host["tet_poses"])
        desc.tet_materials = fptr("tm", host["tet_materials"])
        desc.hex_count = self.hex_count
        desc.hex_indices = iptr("hi", host["hex_indices"])
        desc.hex_inv_J0 = fptr("hj", host["hex_inv_J0"])
        desc.hex_det_J0_w = fptr("hw", host["hex_det_J0_w"])
        desc.hex_materials = fptr("hm", host["hex_materials"])
        desc.fixed_count = int(fixed.size)
        desc.fixed_indices = iptr("fx", fixed)
        desc.spring_count = model.spring_count
        desc.spring_indices = iptr("si", host["spring_indices"])
        desc.spring_rest_length = fptr("sl", host["spring_rest_length"])
        desc.spring_stiffness = fptr("sk", host["spring_stiffness"])
        desc.spring_damping = fptr("sd", host["spring_damping"])
        desc.tri_count = model.tri_count
        desc.tri_indices = iptr("ri", host["tri_indices"])
        desc.tri_poses = fptr("rp", host["tri_poses"])
        desc.tri_materials = fptr("rm", host["tri_materials"])
        if instance_offsets is not None:
            offs = np.asarray(instance_offsets, dtype=np.int32)
            desc.instance_count = int(len(offs) - 1)
            desc.instance_offsets = iptr("io", offs)
            self.instance_count = int(len(offs) - 1)
        else:
            desc.instance_count = 1
            self.instance_count = 1
        handle = ctypes.c_void_p()
        torch.cuda.set_device(self.device.index or 0)
        _lib.check(self._lib.fem_create(ctypes.byref(desc), ctypes.byref(handle)), "fem_create")
        self._sys = handle
        nnz = ctypes.c_int64()
        maxrow = ctypes.c_int32()
        _lib.check(self._lib.fem_pattern_size(self._sys, ctypes.byref(nnz), ctypes.byref(maxrow)), "fem_pattern_size")
        dev = self.device
        self._pc_inv_diag = torch.zeros(max(P, 1), 3, 3, dtype=torch.float32, device=dev)
        self.dv = torch.zeros(max(P, 1), 3, dtype=torch.float32, device=dev)[:P]
        # Matrix storage sized like the reference's A_bsr nnz_capacity, the
        # triplet count P + 2S + 3R + 16T + 64H (solver.py:318-325), so that a
        # topology refresh re-sorts the pattern in place.
        capacity = P + 2 * model.spring_count + 3 * model.tri_count + 16 * model.tet_count + 64 * self.hex_count
        self._bind_matrix(max(capacity, int(nnz.value)))
        self._streaming = False
        self._component_scaling = None
        self._stepped = False
        self.last_update = {}
        self._topology_storage = self._topology_signature()
        _lib.check(self._lib.fem_bind_gravity(self._sys, model.gravity_array.data_ptr()), "fem_bind_gravity")
        if (solver_type, preconditioner_type) != ("cg", "block_diag"):
            _lib.check(self._lib.fem_set_solver(self._sys, SOLVER_TYPES[solver_type],
                                                PRECONDITIONER_TYPES[preconditioner_type], GMRES_RESTART),
                       "fem_set_solver")
        if torque_active:
            _lib.check(self._lib.fem_set_spring_torque(self._sys, fptr("sdir", rest_dir), self.torque_stiffness,
                                                       self.torque_damping), "fem_set_spring_torque")
        if self._particle_contact:
            _lib.check(self._lib.fem_set_particle_contact(
                self._sys, model.particle_ke, model.particle_kd, float(model.particle_kf), float(model.particle_mu),
                float(model.particle_cohesion)), "fem_set_particle_contact")
        self._push_params()
        # Construction-time rest tangent (solver.py:324-341).
        _lib.check(self._lib.fem_assemble(self._sys, model.particle_q.data_ptr(), self.dt, _stream(dev)), "fem_assemble")
        self._anchor = model.particle_q.clone()
        # Per-tet (mu, lambda, k_damp) used by the library; set requires_grad on
        # it to receive material gradients from differentiable steps.
        self.tet_materials = model.tet_materials.detach().clone()
        self._q_asm = model.particle_q.detach().clone() if self.differentiable else None
        self._diff_revision = None
        self._initial_volume = self._compute_volume(model.particle_q)
        self.M_bsr = BlockJacobiOperator(self)

    # ------------------------------------------------------------------
    def _bind_matrix(self, capacity: int):
        """Allocate offsets / columns / values for `capacity` blocks and bind
        them with the inverse diagonal and warm start."""
        P = self.model.particle_count
        dev = self.device
        cap = max(int(capacity), 1)
        self._A_storage = (torch.zeros(P + 1, dtype=torch.int32, device=dev),
                           torch.zeros(cap, dtype=torch.int32, device=dev),
                           torch.zeros(cap, 3, 3, dtype=torch.float32, device=dev))
        offsets, columns, values = self._A_storage
        _lib.check(
            self._lib.fem_bind(self._sys, offsets.data_ptr(), columns.data_ptr(), values.data_ptr(),
                               self._pc_inv_diag.data_ptr(), self.dv.data_ptr(), int(capacity)),
            "fem_bind",
        )
        self._view_matrix()

    def _pattern_nnz(self) -> int:
        nnz = ctypes.c_int64()
        maxrow = ctypes.c_int32()
        _lib.check(self._lib.fem_pattern_size(self._sys, ctypes.byref(nnz), ctypes.byref(maxrow)), "fem_pattern_size")
        return int(nnz.value)

    def memory_stats(self) -> dict:
        """Device allocations the library holds for this solver and their bytes."""
        n, b = ctypes.c_int64(), ctypes.c_int64()
        _lib.check(self._lib.fem_memory_stats(self._sys, ctypes.byref(n), ctypes.byref(b)), "fem_memory_stats")
        return {"allocations": int(n.value), "bytes": int(b.value)}

    def _view_matrix(self):
        offsets, columns, values = self._A_storage
        n = self._pattern_nnz()
        P = self.model.particle_count
        self.A_bsr = BSRMatrix(offsets, columns[:n], values[:n], P, P)

    def _topology_signature(self):
        """Counts and storage of the model arrays a topology refresh reads
        (reference solver.py:859-868)."""
        m = self.model
        counts = (m.particle_count, m.tet_count, self.hex_count, m.spring_count, m.tri_count)
        arrays = (m.particle_mass, m.particle_inv_mass, m.particle_flags, m.tet_indices, m.tet_poses, m.tet_materials)
        return counts, tuple((a.data_ptr(), tuple(a.shape), a.dtype) for a in arrays)

    def prepare_tetrahedral_updates(self, original_node_count: int) -> None:
        """Reserve persistent sparse connections for duplicated Tet4 basis nodes
        (reference solver.py:695-720, StreamingTetAssembly streaming.py:487-597).

        Pool node ``i`` represents original node ``i % original_node_count``;
        every row holds every copy of its original neighbours, so a later cut
        may connect any copies of originally adjacent nodes.  Call before
        capturing or stepping.  Tetrahedra only, no fixed particles.
        """
        if self.fixed_particle_indices.size:
            raise ValueError("Streaming topology is not supported with fixed_particle_indices")
        if self.differentiable:
            raise NotImplementedError("Differentiation requires fixed topology without streaming/component scaling")
        if torch.cuda.is_current_stream_capturing() or self._stepped:
            raise RuntimeError("Prepare tetrahedral updates before capturing or stepping the solver")
        if self._streaming:
            raise RuntimeError("Tetrahedral updates are already prepared")
        if self._component_scaling is not None:
            raise RuntimeError("Prepare tetrahedral updates before component scaling")
        if self.preconditioner_type != "block_diag":
            raise ValueError("Persistent topology updates require the block_diag preconditioner")
        if self.solver_type != "cg":
            raise NotImplementedError("persistent topology updates run solver_type='cg'")
        torch.cuda.synchronize(self.device)
        _lib.check(self._lib.fem_prepare_topology(self._sys, int(original_node_count)), "prepare_tetrahedral_updates")
        self._streaming = True
        self.original_node_count = int(original_node_count)
        self._bind_matrix(self._pattern_nnz())
        st = _stream(self.device)
        _lib.check(self._lib.fem_assemble(self._sys, self.model.particle_q.data_ptr(), self.dt, st), "fem_assemble")
        self._revision += 1

    def refresh_topology(self, particle_q, dt=None, *, element_sources=None):
        """Publish changed tetrahedral connectivity, rest poses, materials and
        masses between graph launches (reference solver.py:870-935).

        Model capacities and array storage must stay unchanged; unused tet
        slots carry zero materials.  After :meth:`prepare_tetrahedral_updates`
        the edit is validated against the reserved adjacency and rejected
        before anything changes, the matrix storage and launch shapes stay
        valid for an existing graph, and :attr:`last_update` reports the
        reference's update counters; ``element_sources`` names each slot's
        previous element (-1: new).  Otherwise the exact pattern is built
        again within the reserved matrix storage, in place when it fits the
        arrays and launches sized at construction (graphs captured before the
        edit then replay the edited topology); otherwise new arrays are made
        and :attr:`topology_reallocated` is set (such graphs must be captured
        again).  Clears the warm start.
        """
        if self.differentiable:
            raise NotImplementedError("Differentiation requires fixed topology without streaming/component scaling")
        dt = self.dt if dt is None else float(dt)
        if torch.cuda.is_current_stream_capturing():
            raise RuntimeError("refresh_topology must run outside CUDA graph capture")
        if dt != self.dt:
            raise ValueError("refresh_topology requires the solver's currently assembled timestep")
        if self._topology_signature() != self._topology_storage:
            raise ValueError("Topology updates must preserve model capacities and array storage")
        P = self.model.particle_count
        _check_vec3("particle_q", particle_q, P, self.device)
        sources = None
        if self._streaming and element_sources is not None:
            if (not isinstance(element_sources, torch.Tensor) or element_sources.dtype != torch.int32
                    or element_sources.dim() != 1 or element_sources.numel() < self.model.tet_count):
                raise ValueError("Element sources must be a capacity-sized int32 array")
            if element_sources.device != self.device:
                raise ValueError("Element sources must reside on the solver device")
            sources = element_sources.contiguous()
        m = self.model
        st = _stream(self.device)
        _lib.check(self._lib.fem_restore_fixed(self._sys, particle_q.data_ptr(), None, st), "fem_restore_fixed")
        stats = (ctypes.c_int32 * 8)()
        _lib.check(
            self._lib.fem_update_topology(
                self._sys, m.tet_indices.data_ptr(), m.tet_poses.data_ptr(), m.tet_materials.data_ptr(),
                m.particle_mass.data_ptr(), m.particle_inv_mass.data_ptr(), m.particle_flags.data_ptr(),
                particle_q.data_ptr(), _ptr(sources), dt, stats, st),
            "refresh_topology",
        )
        if self._streaming:
            keys = ("dirty_elements", "removed_elements", "reused_elements", "active_elements", "touched_rows",
                    "topology_inserted_elements", "topology_removed_elements")
            self.last_update = {k: int(stats[i]) for i, k in enumerate(keys)}
            self.last_update["full_numeric_refresh"] = bool(stats[7])
        else:
            # An edit that did not fit the arrays and launches sized at
            # construction got new ones: graphs captured before it step the
            # old topology and must be captured again.
            self.topology_reallocated = bool(stats[0])
            self._view_matrix()
        self.tet_materials = m.tet_materials.detach().clone()
        self.dv.zero_()
        self._revision += 1

    def prepare_component_scaling(self, node_components: torch.Tensor, component_capacity: int) -> None:
        """Balance disconnected pieces with widely differing masses (reference
        solver.py:722-761, component_scaling.py): every equation of a
        component is divided by the component's total mass, which leaves the
        exact solution and the physical masses unchanged.  ``node_components``
        holds persistent int32 labels (-1 for unused rows); after a topology
        edit update them in place and call :meth:`refresh_component_scaling`.
        """
        if self.fixed_particle_indices.size:
            raise ValueError("Component scaling is not supported with fixed_particle_indices")
        if self.differentiable:
            raise NotImplementedError("Differentiation requires fixed topology without streaming/component scaling")
        if torch.cuda.is_current_stream_capturing() or self._stepped:
            raise RuntimeError("Prepare component scaling before capturing or stepping the solver")
        if self._component_scaling is not None:
            raise RuntimeError("Component scaling is already prepared")
        if self.preconditioner_type != "block_diag":
            raise ValueError("Component scaling requires block_diag preconditioning")
        if self.solver_type != "cg":
            raise NotImplementedError("component scaling runs solver_type='cg'")
        if not isinstance(component_capacity, int) or isinstance(component_capacity, bool) or component_capacity < 1:
            raise ValueError("component_capacity must be a positive integer")
        P = self.model.particle_count
        if (not isinstance(node_components, torch.Tensor) or node_components.dtype != torch.int32
                or tuple(node_components.shape) != (P,) or node_components.device != self.device
                or not node_components.is_contiguous()):
            raise ValueError("node_components must be a particle-sized int32 array on the model device")
        from .component_scaling import ComponentScaling

        self._component_scaling = ComponentScaling(self, node_components, component_capacity)

    def refresh_component_scaling(self) -> None:
        """Recompute the component masses and factors on the device after a
        topology edit (reference solver.py:763-767)."""
        if self._component_scaling is None:
            raise RuntimeError("Prepare component scaling before refreshing it")
        self._component_scaling.refresh()

    def _push_params(self):
        p = _lib.Params()
        p.min_stretch = self.min_stretch
        p.max_stretch = self.max_stretch
        p.has_ground = 0 if self._ground_plane is None else 1
        for k in range(4):
            p.ground[k] = (0.0, 0.0, 1.0, 0.0)[k] if self._ground_plane is None else self._ground_plane[k]
        p.ground_ke, p.ground_kd, p.ground_kf, p.ground_mu = self._ground
        p.linear_damping = self.linear_damping
        for k in range(3):
            p.gravity[k] = self.model.gravity[k]
        p.enable_tet_activations = 1 if self.enable_tet_activations else 0
        p.tolerance = self._tolerance
        p.maxiter = self._maxiter
        p.early_exit = 1 if self._early_exit() else 0
        self._params = p
        self._gravity_pushed = tuple(self.model.gravity)
        _lib.check(self._lib.fem_set_params(self._sys, ctypes.byref(p)), "fem_set_params")

    def _params_ref(self):
        return ctypes.byref(self._params)

    def __del__(self):
        sys = getattr(self, "_sys", None)
        if sys is not None and getattr(self, "_lib", None) is not None:
            try:
                torch.cuda.synchronize(self.device)
                self._lib.fem_destroy(sys)
            except Exception:
                pass
            self._sys = None

    @property
    def linear_solver_tolerance(self) -> float:
        return self._tolerance

    @linear_solver_tolerance.setter
    def linear_solver_tolerance(self, value: float):
        value = float(value)
        if not math.isfinite(value) or value <= 0:
            raise ValueError("linear_solver_tolerance must be finite and positive")
        self._tolerance = value
        self._push_params()

    @property
    def linear_solver_maxiter(self) -> int:
        return self._maxiter

    @linear_solver_maxiter.setter
    def linear_solver_maxiter(self, value: int):
        self._maxiter = int(value)
        self._push_params()

    @property
    def velocity_increment(self) -> torch.Tensor:
        """Persistent solved velocity increment (CG warm start)."""
        return self.dv

    def set_launch(self, elem_lanes: int = 0, assemble_rows: int = 0, cg_workgroups: int = 0):
        """Override kernel launch shapes (0 keeps the current value / auto)."""
        _lib.check(self._lib.fem_set_launch(self._sys, elem_lanes, assemble_rows, cg_workgroups), "fem_set_launch")

    CG_VARIANTS = {"auto": 0, "streamed": 1, "resident": 2}

    def set_cg_variant(self, variant: str = "auto"):
        """Persistent CG kernel: "resident" keeps the matrix on chip for the whole
        solve (one workgroup per CU; raises when a chunk does not fit),
        "streamed" reads it from HBM every iteration, "auto" picks resident
        whenever it fits.  Same iteration either way; only the fp32 summation
        order differs."""
        if variant not in self.CG_VARIANTS:
            raise ValueError(f"cg variant must be one of {sorted(self.CG_VARIANTS)}")
        _lib.check(self._lib.fem_set_cg_variant(self._sys, self.CG_VARIANTS[variant]), "fem_set_cg_variant")

    def launch_config(self) -> dict:
        v = [ctypes.c_int32() for _ in range(6)]
        _lib.check(self._lib.fem_get_launch(self._sys, *[ctypes.byref(x) for x in v]), "fem_get_launch")
        keys = ("elem_lanes", "assemble_rows", "cg_workgroups", "cg_rows_per_thread", "cg_launch_workgroups")
        out = dict(zip(keys, (x.value for x in v[:5])))
        # One persistent CG launch; fewer workgroups than chunks means they
        # take the chunks of whole instances from a device queue.
        out["cg_launch_groups"] = 1
        out["cg_queued"] = out["cg_launch_workgroups"] < out["cg_workgroups"]
        out["cg_variant"] = {1: "streamed", 2: "resident"}[v[5].value]
        return out

    def autotune(self, state: State, dt: float | None = None, *, elem_lanes=(4, 8, 16), assemble_rows=(16, 32, 64),
                 cg_workgroups=None, repeats: int = 3, control: Control | None = None) -> dict:
        """Pick the fastest launch shapes for this model by timing real substeps.

        Each candidate runs ``refresh_tangent`` + ``step`` on copies of
        ``state`` (the caller's state is not modified), starting from the same
        warm start every time; element lanes, assembly rows and the CG
        workgroup count are tuned one after another and the fastest of each
        is kept.  The warm start is restored afterwards.  Returns the chosen
        ``launch_config()``.
        """
        dt = self.dt if dt is None else float(dt)
        P = self.model.particle_count
        dv0 = self.dv.clone()
        s_in, s_out = self.model.state(), self.model.state()
        if cg_workgroups is None:
            base = max(1, P // 160)
            cg_workgroups = sorted({0, 128, 256, 384, 512, base})

        def timed() -> float:
            best = float("inf")
            for r in range(repeats + 1):
                s_in.particle_q.copy_(state.particle_q)
                s_in.particle_qd.copy_(state.particle_qd)
                self.dv.copy_(dv0)
                t0 = torch.cuda.Event(enable_timing=True)
                t1 = torch.cuda.Event(enable_timing=True)
                t0.record()
                self.refresh_tangent(s_in.particle_q, dt)
                self.step(s_in, s_out, control, None, dt)
                t1.record()
                t1.synchronize()
                if r > 0:
                    best = min(best, t0.elapsed_time(t1))
            return best

        for which, cands in ((0, elem_lanes), (1, assemble_rows), (2, cg_workgroups)):
            keep = self.launch_config()
            results = []
            for c in cands:
                args = [0, 0, 0]
                args[which] = int(c)
                try:
                    self.set_launch(*args)
                except (ValueError, _lib.FEMError):
                    continue
                results.append((timed(), c))
            if results:
                _, c = min(results)
                args = [0, 0, 0]
                args[which] = int(c)
                self.set_launch(*args)
            else:
                restore = [keep["elem_lanes"], keep["assemble_rows"], keep["cg_workgroups"]]
                args = [0, 0, 0]
                args[which] = restore[which]
                self.set_launch(*args)
        self.dv.copy_(dv0)
        return self.launch_config()

    # ------------------------------------------------------------------
    def reset(self, state=None, world_mask=None, flags=None):
        """Clear the reusable solve workspace (reference solver.py:436-443)."""
        self.dv.zero_()

    def notify_model_changed(self, flags=None):
        """Re-read gravity from the model (MODEL_PROPERTIES)."""
        self._push_params()

    def refresh_preconditioner(self):
        """The block-Jacobi inverse is rewritten with A by every assembly; nothing to do."""

    def refresh_tangent(self, particle_q: torch.Tensor, dt: float | None = None) -> None:
        """Assemble A = M + h^2 K at ``particle_q`` (reference solver.py:769-857).

        Fixed nodes of ``particle_q`` are first reset to their anchors (as
        the reference does).  A changed timestep clears the warm start.
        """
        dt = self.dt if dt is None else float(dt)
        if not math.isfinite(dt) or dt <= 0:
            raise ValueError("refresh_tangent dt must be finite and positive")
        _check_vec3("particle_q", particle_q, self.model.particle_count, self.device)
        if self._diff is not None and self._diff.recording():
            raise RuntimeError("Reassembly inside a recorded rollout is unsupported; the cached matrix is fixed")
        if dt != self.dt:
            if self.differentiable:
                raise ValueError("Differentiable tangent refresh requires the construction timestep")
            if torch.cuda.is_current_stream_capturing():
                raise RuntimeError("Change the tangent timestep outside tapes and CUDA graph capture")
            if self._streaming:
                raise ValueError("Streaming tangent refresh requires the construction timestep")
            self.dv.zero_()
        st = _stream(self.device)
        _lib.check(self._lib.fem_restore_fixed(self._sys, particle_q.data_ptr(), None, st), "fem_restore_fixed")
        _lib.check(self._lib.fem_assemble(self._sys, particle_q.data_ptr(), float(dt), st), "fem_assemble")
        self.dt = dt
        self._revision += 1
        if self.differentiable:
            self._q_asm = particle_q.detach().clone()

    def rebuild_element_blocks(self, particle_q, dt, *, update_preconditioner=True):
        """Reassemble the deformation-dependent tangent (reference solver.py:506-508)."""
        self.refresh_tangent(particle_q, dt)

    def _early_exit(self) -> bool:
        # Device-side residual exit: the reference stops early with the CUDA
        # graph's conditional loop (linear_solver_use_cuda_graph) or with host
        # checks every linear_solver_check_every iterations (solver.py:1203-1217);
        # otherwise it runs the full budget with residual-masked updates.
        return self.linear_solver_use_cuda_graph or self.linear_solver_check_every > 0

    def _sync_params(self):
        if tuple(self.model.gravity) != self._gravity_pushed or (self._params.early_exit != 0) != self._early_exit():
            self._push_params()

    def step(self, state_in: State, state_out: State, control: Control | None, contacts, dt: float, *,
             external_forces: torch.Tensor | None = None, observables=None) -> State:
        """One linearly implicit substep (reference solver.py:942-1002).

        ``state_in.particle_f`` is overwritten with the projected right-hand
        side; ``state_in`` positions/velocities of fixed nodes are reset to
        their anchors; ``state_out`` receives the integrated state.
        """
        P = self.model.particle_count
        if external_forces is not None:
            _check_vec3("external_forces", external_forces, P, self.device)
        if not (math.isfinite(dt) and dt > 0):
            raise ValueError("dt must be finite and positive")
        if self.differentiable:
            from .differentiation import differentiable_step

            return differentiable_step(self, state_in, state_out, control, contacts, dt, external_forces)
        if torch.is_grad_enabled() and any(
                x is not None and x.requires_grad
                for x in (state_in.particle_q, state_in.particle_qd, external_forces)):
            # The reference refuses a step of a non-differentiable solver on a
            # recording tape (solver.py:985-986); inputs that require
            # gradients are the autograd counterpart.
            raise RuntimeError("Construct the solver with differentiable=True before recording a step")
        self._bind_soft_contacts(contacts)
        for name, t in (("state_in.particle_q", state_in.particle_q), ("state_in.particle_qd", state_in.particle_qd),
                        ("state_in.particle_f", state_in.particle_f), ("state_out.particle_q", state_out.particle_q),
                        ("state_out.particle_qd", state_out.particle_qd)):
            _check_vec3(name, t, P, self.device)
        self._sync_params()
        act = None
        if self.enable_tet_activations:
            ctrl = control if control is not None else self.model.control()
            act = ctrl.tet_activations
            if act.dtype != torch.float32 or act.numel() != self.model.tet_count or act.device != self.device:
                raise ValueError("control.tet_activations must be a float32 (tet_count,) tensor on the model device")
            act = act.contiguous()
        self._stepped = True
        st = _stream(self.device)
        _lib.check(
            self._lib.fem_step(self._sys, state_in.particle_q.data_ptr(), state_in.particle_qd.data_ptr(),
                               state_in.particle_f.data_ptr(), _ptr(external_forces), _ptr(act),
                               state_out.particle_q.data_ptr(), state_out.particle_qd.data_ptr(), float(dt), st),
            "fem_step",
        )
        if self.linear_solver_check_every > 0:
            # Host-check path (solver.py:1333-1335): the reference tests the
            # residual every check_every iterations, so it reports the first
            # multiple of check_every at or after convergence (capped at maxiter).
            it = int(self.solve_stats()["iterations"][0])
            n = int(self.linear_solver_check_every)
            self.last_cg_iters = min(self.linear_solver_maxiter, -(-it // n) * n)
        return state_out

    def implicit_integration(self, model, state_in: State, state_out: State, dt: float) -> None:
        """Solve ``A dv = h f - h^2 K v`` and integrate ``v``, ``x`` (reference
        solver.py:1254-1351), the second half of :meth:`step`.

        ``state_in.particle_f`` must hold the force sum already multiplied by
        ``dt`` (external loads included); it is overwritten with the projected
        right-hand side.  The input state's fixed nodes are not reset (the
        reference's step does that before the forces); ``state_out`` receives
        the integrated state with fixed nodes at their anchors.
        """
        P = self.model.particle_count
        if not (math.isfinite(dt) and dt > 0):
            raise ValueError("dt must be finite and positive")
        for name, t in (("state_in.particle_q", state_in.particle_q), ("state_in.particle_qd", state_in.particle_qd),
                        ("state_in.particle_f", state_in.particle_f), ("state_out.particle_q", state_out.particle_q),
                        ("state_out.particle_qd", state_out.particle_qd)):
            _check_vec3(name, t, P, self.device)
        self._sync_params()
        self._stepped = True
        st = _stream(self.device)
        _lib.check(self._lib.fem_integrate_rhs(self._sys, state_in.particle_qd.data_ptr(),
                                               state_in.particle_f.data_ptr(), float(dt), st), "fem_integrate_rhs")
        _lib.check(
            self._lib.fem_step_solve(self._sys, state_in.particle_q.data_ptr(), state_in.particle_qd.data_ptr(),
                                     state_in.particle_f.data_ptr(), state_out.particle_q.data_ptr(),
                                     state_out.particle_qd.data_ptr(), float(dt), st),
            "fem_step_solve",
        )
        if self.linear_solver_check_every > 0:
            it = int(self.solve_stats()["iterations"][0])
            n = int(self.linear_solver_check_every)
            self.last_cg_iters = min(self.linear_solver_maxiter, -(-it // n) * n)

    def prepare_step(self, state_in: State, control: Control | None, contacts, dt: float, *,
                     external_forces: torch.Tensor | None = None):
        """Prepare reusable affine mechanics without integrating physical state
        (reference solver.py:483-500).  Returns the solver's persistent
        :class:`PreparedFEMStep`; the caller owns the tangent refresh and must
        prepare again for every new accepted state, control or matrix revision.
        """
        from .prepared_step import PreparedFEMStep

        PreparedFEMStep.check_supported(self)
        if torch.cuda.is_current_stream_capturing():
            raise RuntimeError("Prepare FEM mechanics before CUDA graph capture")
        self._bind_soft_contacts(contacts)
        if not hasattr(self, "_prepared_step"):
            self._prepared_step = PreparedFEMStep(self)
        return self._prepared_step.prepare(state_in, control, contacts, dt, external_forces)

    def _bind_soft_contacts(self, contacts) -> None:
        """Bind the contact list's soft contacts (device arrays, read at run
        time) for the next force evaluations, or unbind them when `contacts`
        has none (the reference's eval_soft_contact_forces returns zeros then,
        solver.py:1543-1544).  Rebinding different arrays inside a CUDA graph
        capture is refused: a recorded graph keeps the arrays of its capture."""
        if contacts is None or not hasattr(contacts, "soft_contact_count"):
            key = None
            arrays = None
        else:
            dev = self.device

            def arr(name, dtype, shape):
                x = torch.as_tensor(getattr(contacts, name), device=dev)
                if x.dtype != dtype or tuple(x.shape) != shape or not x.is_contiguous():
                    raise ValueError(f"contacts.{name} must be a contiguous {dtype} {shape} array on the model device")
                return x

            count = torch.as_tensor(contacts.soft_contact_count, device=dev)
            if count.dtype != torch.int32 or count.numel() < 1:
                raise ValueError("contacts.soft_contact_count must be an int32 device array holding the count")
            part = torch.as_tensor(contacts.soft_contact_particle, device=dev)
            m = int(getattr(contacts, "soft_contact_max", part.numel()))
            if m < 0 or m > part.numel():
                raise ValueError("contacts.soft_contact_max exceeds the contact arrays")
            arrays = (count, arr("soft_contact_particle", torch.int32, (part.numel(),)),
                      arr("soft_contact_body_pos", torch.float32, (part.numel(), 3)),
                      arr("soft_contact_body_vel", torch.float32, (part.numel(), 3)),
                      arr("soft_contact_normal", torch.float32, (part.numel(), 3)))
            mdl = self.model
            coef = (float(mdl.soft_contact_ke), float(mdl.soft_contact_kd), float(mdl.soft_contact_kf),
                    float(mdl.soft_contact_mu))
            key = tuple(a.data_ptr() for a in arrays) + (m,) + coef
        if key == getattr(self, "_soft_key", None):
            return
        if torch.cuda.is_current_stream_capturing():
            raise RuntimeError("Step once with these contacts before CUDA graph capture")
        if key is None:
            _lib.check(self._lib.fem_set_soft_contacts(self._sys, None, None, None, None, None, 0, 0.0, 0.0, 0.0, 0.0),
                       "fem_set_soft_contacts")
        else:
            _lib.check(self._lib.fem_set_soft_contacts(self._sys, *(a.data_ptr() for a in arrays), m, *coef),
                       "fem_set_soft_contacts")
        self._soft_key = key
        self._soft_arrays = arrays  # the library reads them until the next binding

    def eval_soft_contact_forces(self, model, state, contacts):
        """Particle-shape penalty + damping + Coulomb friction [N] of the
        contact list (reference solver.py:1540-1566; zero without one)."""
        self._bind_soft_contacts(contacts)
        if self._soft_key is None:
            return torch.zeros_like(state.particle_q)
        return self._eval(_lib.FEM_TERM_SOFT, state)

    def set_barrier_spin_limit(self, spins: int = 0) -> None:
        """Bound the persistent CG's cross-workgroup waits (flag polls; 0 =
        default).  A wait that runs out marks the solve as timed out and the
        step writes NaN instead of a partial result (see :meth:`solve_stats`)."""
        _lib.check(self._lib.fem_set_barrier_spin_limit(self._sys, int(spins)), "set_barrier_spin_limit")

    def solve_stats(self) -> dict:
        """Per-instance statistics of the last solve (synchronizes)."""
        n = self.instance_count
        it = (ctypes.c_int32 * n)()
        rr = (ctypes.c_float * n)()
        at = (ctypes.c_float * n)()
        er = (ctypes.c_int32 * n)()
        _lib.check(self._lib.fem_solve_stats(self._sys, _stream(self.device), it, rr, at, er), "fem_solve_stats")
        if any(er):
            raise _lib.FEMError("CG barrier timed out: the persistent launch was not fully resident")
        return {"iterations": list(it), "residual_sq": list(rr), "atol_sq": list(at)}

    # ------------------------------------------------------------------
    # Force evaluators (reference solver.py:1357-1556), unscaled [N].
    def _eval(self, terms: int, state: State, control: Control | None = None) -> torch.Tensor:
        self._sync_params()
        out = torch.empty_like(state.particle_q)
        act = None
        if self.enable_tet_activations and (terms & _lib.FEM_TERM_TET):
            act = (control if control is not None else self.model.control()).tet_activations.contiguous()
        _lib.check(
            self._lib.fem_eval_forces(self._sys, terms, state.particle_q.data_ptr(), state.particle_qd.data_ptr(),
                                      _ptr(act), out.data_ptr(), _stream(self.device)),
            "fem_eval_forces",
        )
        return out

    def eval_spring_forces(self, model, state):
        return self._eval(_lib.FEM_TERM_SPRING, state)

    def eval_triangle_forces(self, model, control, state):
        return self._eval(_lib.FEM_TERM_TRI, state)

    def eval_tetrahedral_forces(self, model, control, state):
        return self._eval(_lib.FEM_TERM_TET, state, control)

    def eval_hexahedral_forces(self, model, control, state):
        return self._eval(_lib.FEM_TERM_HEX, state, control)

    def eval_particle_particle_forces(self, model, control, state):
        """Hash-grid particle-particle penalty + Coulomb friction [N] (zero
        when the model has no particle grid, reference solver.py:1477-1510)."""
        if not self._particle_contact:
            return torch.zeros_like(state.particle_q)
        return self._eval(_lib.FEM_TERM_PARTICLE, state)

    def eval_particle_ground_contact_forces(self, model, control, state):
        return self._eval(_lib.FEM_TERM_GROUND, state)

    def eval_gravity_forces(self, model, state=None):
        s = state if state is not None else State(model.particle_q, model.particle_qd, None)
        return self._eval(_lib.FEM_TERM_GRAVITY, s)

    def eval_linear_damping_forces(self, model, state):
        return self._eval(_lib.FEM_TERM_DAMPING, state)

    # ------------------------------------------------------------------
    # Volume telemetry (reference solver.py:1614-1701).
    def _compute_volume(self, positions: torch.Tensor | None) -> float:
        total = 0.0
        m = self.model
        if m.tet_count and positions is not None:
            idx = m.tet_indices.long()
            p = positions[idx]
            e = p[:, 1:] - p[:, :1]
            vol = torch.linalg.det(e.double()).abs() / 6.0
            if self._tet_volume_weights is not None:
                vol = vol * self._tet_volume_weights.double()
            total += float(vol.sum())
        if self.hex_count:
            total += float(m.fem.hex_det_J0_w.double().sum())
        return total

    def compute_volume(self, state: State | None = None) -> float:
        return self._compute_volume(None if state is None else state.particle_q)

    def get_initial_volume(self, state=None) -> float:
        return self._initial_volume if self._initial_volume > 1e-12 else 1.0

    def get_volume_ratio(self, state=None) -> float:
        initial = self.get_initial_volume(state)
        current = self.compute_volume(state)
        if not math.isfinite(current) or current <= 0.0:
            return 1.0
        ratio = current / initial
        return ratio if math.isfinite(ratio) else 1.0

    def set_tet_materials(self, materials: torch.Tensor) -> None:
        """Replace the per-tet (mu, lambda, k_damp) and re-assemble at the last assembly positions.

        No reference counterpart (there the materials are fixed model inputs);
        for material optimisation with ``differentiable=True``.  Tets keep their
        zero / non-zero status (zero-material tets are not in the pattern).
        """
        m = torch.as_tensor(materials).reshape(-1, 3)
        if m.shape[0] != self.model.tet_count or m.dtype != torch.float32 or m.device != self.device:
            raise ValueError("materials must be a float32 (tet_count, 3) tensor on the solver's device")
        # Kept by reference: when it requires a gradient (e.g. it is a function
        # of optimisation parameters), differentiable steps propagate into it.
        self.tet_materials = m
        self._tet_materials_values = m.detach().contiguous()
        st = _stream(self.device)
        _lib.check(self._lib.fem_set_tet_materials(self._sys, self._tet_materials_values.data_ptr(), st),
                   "fem_set_tet_materials")
        q = self._q_asm if self._q_asm is not None else self.model.particle_q
        _lib.check(self._lib.fem_assemble(self._sys, q.data_ptr(), float(self.dt), st), "fem_assemble")
        self._revision += 1

    # Differentiation entry points (reference solver.py:445-469).
    def prepare_differentiation(self):
        if not self.differentiable:
            raise RuntimeError("Construct the solver with differentiable=True")
        from .differentiation import prepare

        return prepare(self)

    def check_differentiation(self, *, reset=True):
        if not self.differentiable:
            raise RuntimeError("Construct the solver with differentiable=True")
        from .differentiation import check

        return check(self, reset=reset)


SolverLinearFEM = SolverFEM
