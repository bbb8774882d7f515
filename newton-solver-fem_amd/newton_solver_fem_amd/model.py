"""Minimal model/state containers and builders over torch device tensors.

The reference builds its models with Newton 1.7's ``ModelBuilder`` plus the
package's ``FEMBuilder`` (``src/newton_solver_fem/_src/solvers/fem/builder.py``).
Newton is not a dependency here, so this module supplies the subset of that
surface the FEM substep consumes: particles with mass/inverse mass/radius/
flags, Tet4 elements with inverted rest matrices (``tet_poses``) and (mu,
lambda, k_damp) materials, linear springs, triangle membranes with inverted
2x2 rest matrices (``tri_poses``), Hex8 reference quadrature data under
``model.fem``, gravity, and ``state()`` / ``control()`` factories.

Host copies of every array are kept on the model (``model.host``) because the
sparsity pattern and element incidence lists are built on the host once.
"""

from __future__ import annotations

import warnings
from dataclasses import dataclass

import numpy as np
import torch

PARTICLE_FLAG_ACTIVE = 1

# Hex8 corner / Gauss-point sign table (reference builder.py:104-107).
HEX_SIGNS = np.asarray(
    [[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1], [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]],
    dtype=np.float64,
)


def hex_reference_data(rest: np.ndarray, hexes: np.ndarray):
    """Per hex and Gauss point: ``inv(J0)`` and ``|det J0|`` in FP64.

    Restates ``FEMBuilder.finalize`` (reference builder.py:95-135):
    ``J0 = sum_a x_a (outer) grad_xi N_a(xi_g)`` at the 2x2x2 Gauss points
    ``signs/sqrt(3)`` with unit Gauss weights; raises if ``det J0 <= 1e-12``.
    """
    rest = np.asarray(rest, dtype=np.float64)
    hexes = np.asarray(hexes, dtype=np.int64).reshape(-1, 8)
    gp = HEX_SIGNS / np.sqrt(3.0)  # (8 points, 3)
    s = HEX_SIGNS  # (8 corners, 3)
    xi, eta, zeta = gp[:, 0:1], gp[:, 1:2], gp[:, 2:3]  # (8,1) each
    xa, ya, za = s[None, :, 0], s[None, :, 1], s[None, :, 2]  # (1,8)
    grad = np.stack(
        [
            xa * (1 + ya * eta) * (1 + za * zeta),
            (1 + xa * xi) * ya * (1 + za * zeta),
            (1 + xa * xi) * (1 + ya * eta) * za,
        ],
        axis=-1,
    ) * 0.125  # (8 points, 8 corners, 3)
    pts = rest[hexes]  # (H, 8 corners, 3)
    jac = np.einsum("hai,gaj->hgij", pts, grad)  # (H, 8 points, 3, 3)
    det = np.linalg.det(jac)
    bad = np.argwhere(det <= 1.0e-12)
    if len(bad):
        e, g = bad[0]
        raise ValueError(f"Degenerate or inverted Hex8 {e}, Gauss point {g}: {det[e, g]}")
    return np.linalg.inv(jac), np.abs(det)


def _rigid(xform):
    """Rotation matrix and translation of a ``(px, py, pz, qx, qy, qz, qw)`` transform."""
    v = np.asarray(xform, dtype=np.float64).reshape(7)
    p = v[:3]
    x, y, z, w = v[3:] / np.linalg.norm(v[3:])
    R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                  [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                  [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
    return R, p


@dataclass
class LinearFEMData:
    """Hex8 reference geometry and materials stored under ``model.fem``."""

    hex_count: int
    hex_indices: torch.Tensor
    hex_materials: torch.Tensor
    hex_inv_J0: torch.Tensor
    hex_det_J0_w: torch.Tensor


class State:
    """Particle positions, velocities and the force workspace."""

    def __init__(self, particle_q: torch.Tensor, particle_qd: torch.Tensor, particle_f: torch.Tensor):
        self.particle_q = particle_q
        self.particle_qd = particle_qd
        self.particle_f = particle_f

    def clear_forces(self):
        self.particle_f.zero_()


class Control:
    def __init__(self, tet_activations: torch.Tensor):
        self.tet_activations = tet_activations


class Contacts:
    """The soft-contact part of Newton's ``Contacts`` (the collision
    pipeline's particle-shape contact list, read by the reference's
    eval_soft_contact_forces solver.py:1540-1566): ``soft_contact_count`` an
    int32 (1,) device tensor with the number of valid entries,
    ``soft_contact_particle`` int32 (max,), ``soft_contact_body_pos``,
    ``soft_contact_body_vel`` and ``soft_contact_normal`` float32 (max, 3)
    (the normal points from the body to the particle), and
    ``soft_contact_max`` = max.  Any object with these attributes (torch
    tensors, or arrays torch.as_tensor takes without a copy) works as well."""

    def __init__(self, max_contacts: int, device):
        m = int(max_contacts)
        if m < 0:
            raise ValueError("max_contacts must be nonnegative")
        self.soft_contact_max = m
        self.soft_contact_count = torch.zeros(1, dtype=torch.int32, device=device)
        self.soft_contact_particle = torch.full((m,), -1, dtype=torch.int32, device=device)
        self.soft_contact_body_pos = torch.zeros((m, 3), dtype=torch.float32, device=device)
        self.soft_contact_body_vel = torch.zeros((m, 3), dtype=torch.float32, device=device)
        self.soft_contact_normal = torch.zeros((m, 3), dtype=torch.float32, device=device)


class HashGrid:
    """Marks a model whose particles collide with each other (Newton's
    ``Model.particle_grid``, a Warp hash grid).  The device grid itself is
    owned and rebuilt every step by the solver library."""


class Model:
    """Finalized arrays on one torch device (``cuda`` for the HIP solver)."""

    def __init__(self, device, host: dict, gravity):
        self.device = torch.device(device)
        self.host = host
        # Device copy of gravity that solvers read at run time.
        self.gravity_array = torch.zeros(3, dtype=torch.float32, device=self.device)
        self.gravity = gravity
        self.particle_count = int(len(host["particle_q"]))
        # Hash grid of particle-particle contact.  None (the default of
        # finalize) skips that force term, as the reference does when
        # Model.particle_grid is None (solver.py:1477-1485,
        # kernels.py:1032-1038); finalize(particle_grid=True) sets a
        # HashGrid and SolverFEM then evaluates the term.  Contact
        # coefficients follow Newton's Model fields; SolverFEM writes
        # particle_ke / particle_kd from self_contact_ke / self_contact_kd.
        self.particle_grid = None
        self.particle_ke = 1.0e3
        self.particle_kd = 1.0e2
        self.particle_kf = 1.0e2
        self.particle_mu = 0.5
        self.particle_cohesion = 0.0
        # Particle-shape soft contact coefficients (Newton's Model
        # soft_contact_ke / kd / kf / mu, read by the reference's
        # eval_soft_contact_forces solver.py:1540-1566; the values are
        # Newton's ModelBuilder defaults).
        self.soft_contact_ke = 1.0e3
        self.soft_contact_kd = 10.0
        self.soft_contact_kf = 1.0e3
        self.soft_contact_mu = 0.5
        r = np.asarray(host["particle_radius"], dtype=np.float64)
        r = r[np.isfinite(r)]
        self.particle_max_radius = float(r.max()) if r.size else 0.0
        self.tet_count = int(len(host["tet_indices"]))
        self.spring_count = int(len(host["spring_indices"]))
        self.tri_count = int(len(host["tri_indices"]))
        dev = self.device

        def t(a, dtype):
            return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype, device=dev)

        self.particle_q = t(host["particle_q"], torch.float32).reshape(-1, 3)
        self.particle_qd = t(host["particle_qd"], torch.float32).reshape(-1, 3)
        self.particle_mass = t(host["particle_mass"], torch.float32)
        self.particle_inv_mass = t(host["particle_inv_mass"], torch.float32)
        self.particle_radius = t(host["particle_radius"], torch.float32)
        self.particle_flags = t(host["particle_flags"], torch.int32)
        self.tet_indices = t(host["tet_indices"], torch.int32).reshape(-1, 4)
        self.tet_poses = t(host["tet_poses"], torch.float32).reshape(-1, 3, 3)
        self.tet_materials = t(host["tet_materials"], torch.float32).reshape(-1, 3)
        self.tet_activations = t(host["tet_activations"], torch.float32)
        self.spring_indices = t(host["spring_indices"], torch.int32).reshape(-1)
        self.spring_rest_length = t(host["spring_rest_length"], torch.float32)
        self.spring_stiffness = t(host["spring_stiffness"], torch.float32)
        self.spring_damping = t(host["spring_damping"], torch.float32)
        self.tri_indices = t(host["tri_indices"], torch.int32).reshape(-1, 3)
        self.tri_poses = t(host["tri_poses"], torch.float32).reshape(-1, 2, 2)
        self.tri_materials = t(host["tri_materials"], torch.float32).reshape(-1, 5)
        self.tri_areas = t(host["tri_areas"], torch.float32)
        h = host["hex_indices"]
        self.fem = LinearFEMData(
            len(h),
            t(h, torch.int32).reshape(-1, 8),
            t(host["hex_materials"], torch.float32).reshape(-1, 3),
            t(host["hex_inv_J0"], torch.float32).reshape(-1, 8, 3, 3),
            t(host["hex_det_J0_w"], torch.float32).reshape(-1, 8),
        )

    @property
    def gravity(self) -> tuple:
        return self._gravity

    @gravity.setter
    def gravity(self, value):
        """Setting gravity updates ``gravity_array`` in place: solvers read it on
        the device, so graphs recorded earlier use the new value (the
        reference's Model.gravity is a device array)."""
        self._gravity = tuple(float(g) for g in value)
        self.gravity_array.copy_(torch.tensor(self._gravity, dtype=torch.float32))

    def set_gravity(self, gravity):
        self.gravity = gravity

    def state(self, requires_grad: bool = False) -> State:
        q = self.particle_q.clone()
        qd = self.particle_qd.clone()
        f = torch.zeros_like(q)
        if requires_grad:
            q.requires_grad_(True)
            qd.requires_grad_(True)
        return State(q, qd, f)

    def control(self) -> Control:
        return Control(self.tet_activations.clone())


class ModelBuilder:
    """Collects particles, Tet4 elements, springs and triangles (Newton ``ModelBuilder`` subset)."""

    default_particle_radius = 0.1
    # Newton's ModelBuilder defaults for triangle membranes.
    default_tri_ke = 100.0
    default_tri_ka = 100.0
    default_tri_kd = 10.0
    default_tri_drag = 0.0
    default_tri_lift = 0.0

    def __init__(self, gravity=-9.81, up_axis: int = 2):
        if np.ndim(gravity) == 0:
            g = [0.0, 0.0, 0.0]
            g[up_axis] = float(gravity)
            gravity = g
        self.gravity = tuple(float(x) for x in gravity)
        self._q, self._qd, self._m, self._r, self._flags = [], [], [], [], []
        self._tet_idx, self._tet_pose, self._tet_mat, self._tet_act = [], [], [], []
        self._spring_idx, self._spring_rest, self._spring_ke, self._spring_kd, self._spring_control = [], [], [], [], []
        self._tri_idx, self._tri_pose, self._tri_mat, self._tri_area = [], [], [], []

    @property
    def particle_count(self) -> int:
        return len(self._m)

    @property
    def particle_q(self):
        return list(self._q)

    @property
    def particle_mass(self):
        """The builder's particle masses, as Newton's mutable list (assigning an
        entry, e.g. 0 for a support, changes the particle's mass)."""
        return self._m

    @property
    def tet_count(self) -> int:
        return sum(len(a) for a in self._tet_idx)

    @property
    def spring_count(self) -> int:
        return len(self._spring_ke)

    @property
    def spring_stiffness(self):
        return list(self._spring_ke)

    @property
    def spring_damping(self):
        return list(self._spring_kd)

    @property
    def tri_count(self) -> int:
        return len(self._tri_area)

    def add_spring(self, i: int, j: int, ke: float, kd: float, control: float = 0.0) -> None:
        """Linear spring between particles i and j; the rest length is their current distance."""
        self.add_springs([(i, j)], ke, kd, control)

    def add_springs(self, pairs, ke, kd, control=0.0) -> None:
        """Vectorized ``add_spring`` (rest lengths from the current positions, FP64)."""
        pairs = np.asarray(pairs, dtype=np.int64).reshape(-1, 2)
        q = np.asarray(self._q, dtype=np.float64)
        rest = np.linalg.norm(q[pairs[:, 0]] - q[pairs[:, 1]], axis=1)
        n = len(pairs)
        self._spring_idx.extend(map(tuple, pairs.tolist()))
        self._spring_rest.extend(rest.tolist())
        self._spring_ke.extend(np.broadcast_to(np.asarray(ke, dtype=np.float64), (n,)).tolist())
        self._spring_kd.extend(np.broadcast_to(np.asarray(kd, dtype=np.float64), (n,)).tolist())
        self._spring_control.extend(np.broadcast_to(np.asarray(control, dtype=np.float64), (n,)).tolist())

    def add_triangle(self, i: int, j: int, k: int, tri_ke=None, tri_ka=None, tri_kd=None, tri_drag=None,
                     tri_lift=None) -> float:
        """Triangle membrane; returns its rest area (0 and nothing added when degenerate).

        The rest pose is the inverse of the 2x2 matrix of the edges (x_j - x_i,
        x_k - x_i) in an in-plane basis (e1 along the first edge, e2 = n x e1).
        """
        return float(self.add_triangles([(i, j, k)], tri_ke, tri_ka, tri_kd, tri_drag, tri_lift)[0])

    def add_triangles(self, indices, tri_ke=None, tri_ka=None, tri_kd=None, tri_drag=None, tri_lift=None):
        """Vectorized ``add_triangle``; returns the rest areas (0 for skipped triangles)."""
        idx = np.asarray(indices, dtype=np.int64).reshape(-1, 3)
        q = np.asarray(self._q, dtype=np.float64)
        qp, rp = q[idx[:, 1]] - q[idx[:, 0]], q[idx[:, 2]] - q[idx[:, 0]]
        n = np.cross(qp, rp)
        nl, el = np.linalg.norm(n, axis=1), np.linalg.norm(qp, axis=1)
        area = np.zeros(len(idx))
        ok = (nl > 0.0) & (el > 0.0)
        e1 = qp[ok] / el[ok, None]
        e2 = np.cross(n[ok] / nl[ok, None], e1)
        e2 /= np.linalg.norm(e2, axis=1)[:, None]
        D = np.empty((int(ok.sum()), 2, 2))
        D[:, 0, 0], D[:, 0, 1] = np.sum(e1 * qp[ok], 1), np.sum(e1 * rp[ok], 1)
        D[:, 1, 0], D[:, 1, 1] = np.sum(e2 * qp[ok], 1), np.sum(e2 * rp[ok], 1)
        area[ok] = np.linalg.det(D) / 2.0
        keep = area > 0.0
        if not keep.all():
            warnings.warn(f"skipping {int((~keep).sum())} inverted or degenerate triangle element(s)")
        Dk = D[keep[ok]]
        d = lambda v, dflt: float(dflt if v is None else v)
        mat = (d(tri_ke, self.default_tri_ke), d(tri_ka, self.default_tri_ka), d(tri_kd, self.default_tri_kd),
               d(tri_drag, self.default_tri_drag), d(tri_lift, self.default_tri_lift))
        self._tri_idx.extend(map(tuple, idx[keep].tolist()))
        self._tri_pose.extend(np.linalg.inv(Dk) if len(Dk) else [])
        self._tri_mat.extend([mat] * int(keep.sum()))
        self._tri_area.extend(area[keep].tolist())
        return area

    def add_particle(self, pos, vel=(0.0, 0.0, 0.0), mass=1.0, radius=None, flags=PARTICLE_FLAG_ACTIVE) -> int:
        self._q.append(np.asarray(pos, dtype=np.float64).reshape(3))
        self._qd.append(np.asarray(vel, dtype=np.float64).reshape(3))
        self._m.append(float(mass))
        self._r.append(self.default_particle_radius if radius is None else float(radius))
        self._flags.append(int(flags))
        return len(self._m) - 1

    def add_particles(self, pos, vel, mass, radius=None):
        pos = np.asarray(pos, dtype=np.float64).reshape(-1, 3)
        vel = np.asarray(vel, dtype=np.float64).reshape(-1, 3)
        mass = np.broadcast_to(np.asarray(mass, dtype=np.float64), (len(pos),))
        rad = np.broadcast_to(
            np.asarray(self.default_particle_radius if radius is None else radius, dtype=np.float64), (len(pos),)
        )
        self._q.extend(pos)
        self._qd.extend(vel)
        self._m.extend(mass.tolist())
        self._r.extend(rad.tolist())
        self._flags.extend([PARTICLE_FLAG_ACTIVE] * len(pos))

    def add_tetrahedron(self, i, j, k, l, k_mu=1.0e3, k_lambda=1.0e3, k_damp=0.0) -> float:
        """Append one tet; its pose is ``inv([x1-x0, x2-x0, x3-x0])`` in FP64.

        Like Newton's builder, a non-positive rest volume skips the element.
        """
        return float(self.add_tetrahedra(np.array([[i, j, k, l]]), [[k_mu, k_lambda, k_damp]])[0])

    def add_tetrahedra(self, indices, materials, activations=None) -> np.ndarray:
        """Vectorized ``add_tetrahedron``; returns the rest volumes."""
        idx = np.asarray(indices, dtype=np.int64).reshape(-1, 4)
        mats = np.broadcast_to(np.asarray(materials, dtype=np.float64), (len(idx), 3))
        pts = np.asarray(self._q, dtype=np.float64)
        dm = np.transpose(pts[idx[:, 1:]] - pts[idx[:, :1]], (0, 2, 1))  # columns = edges
        det = np.linalg.det(dm)
        volume = det / 6.0
        ok = volume > 0.0
        if not ok.all():
            warnings.warn(f"skipping {int((~ok).sum())} inverted tetrahedral element(s)")
        poses = np.linalg.inv(dm[ok])
        self._tet_idx.append(idx[ok].astype(np.int32))
        self._tet_pose.append(poses)
        self._tet_mat.append(np.array(mats[ok]))
        act = np.zeros(int(ok.sum())) if activations is None else np.asarray(activations, dtype=np.float64)[ok]
        self._tet_act.append(act)
        return volume

    def add_builder(self, builder: "ModelBuilder", xform=None, label_prefix=None) -> None:
        """Append another builder's particles and elements (Newton's
        ``ModelBuilder.add_builder``).  ``xform`` is a rigid transform
        ``(px, py, pz, qx, qy, qz, qw)`` (a ``wp.transform``'s layout) applied to
        the copied positions and, by its rotation, velocities; element rest
        data (tet and triangle poses, spring rest lengths) are invariant under
        it and copied as they are.  ``label_prefix`` is accepted for Newton
        compatibility (this builder keeps no labels)."""
        offset = self.particle_count
        q = np.asarray(builder._q, dtype=np.float64).reshape(-1, 3)
        qd = np.asarray(builder._qd, dtype=np.float64).reshape(-1, 3)
        if xform is not None:
            R, p = _rigid(xform)
            q = q @ R.T + p
            qd = qd @ R.T
        self._q.extend(q)
        self._qd.extend(qd)
        self._m.extend(builder._m)
        self._r.extend(builder._r)
        self._flags.extend(builder._flags)
        for idx, pose, mat, act in zip(builder._tet_idx, builder._tet_pose, builder._tet_mat, builder._tet_act):
            self._tet_idx.append((np.asarray(idx) + offset).astype(np.int32))
            self._tet_pose.append(np.array(pose))
            self._tet_mat.append(np.array(mat))
            self._tet_act.append(np.array(act))
        self._spring_idx.extend((i + offset, j + offset) for i, j in builder._spring_idx)
        self._spring_rest.extend(builder._spring_rest)
        self._spring_ke.extend(builder._spring_ke)
        self._spring_kd.extend(builder._spring_kd)
        self._spring_control.extend(builder._spring_control)
        self._tri_idx.extend((i + offset, j + offset, k + offset) for i, j, k in builder._tri_idx)
        self._tri_pose.extend(np.array(x) for x in builder._tri_pose)
        self._tri_mat.extend(builder._tri_mat)
        self._tri_area.extend(builder._tri_area)

    def _host_arrays(self) -> dict:
        p = len(self._m)
        mass = np.asarray(self._m, dtype=np.float64)
        inv_mass = np.where(mass > 0.0, 1.0 / np.where(mass > 0.0, mass, 1.0), 0.0)
        cat = (lambda lst, shape, dt: np.concatenate(lst).astype(dt) if lst else np.zeros(shape, dt))
        return {
            "particle_q": np.asarray(self._q, dtype=np.float32).reshape(p, 3),
            "particle_qd": np.asarray(self._qd, dtype=np.float32).reshape(p, 3),
            "particle_mass": mass.astype(np.float32),
            "particle_inv_mass": inv_mass.astype(np.float32),
            "particle_radius": np.asarray(self._r, dtype=np.float32),
            "particle_flags": np.asarray(self._flags, dtype=np.int32),
            "tet_indices": cat(self._tet_idx, (0, 4), np.int32),
            "tet_poses": cat(self._tet_pose, (0, 3, 3), np.float32),
            "tet_materials": cat(self._tet_mat, (0, 3), np.float32),
            "tet_activations": cat(self._tet_act, (0,), np.float32),
            "spring_indices": np.asarray(self._spring_idx, dtype=np.int32).reshape(-1, 2),
            "spring_rest_length": np.asarray(self._spring_rest, dtype=np.float32),
            "spring_stiffness": np.asarray(self._spring_ke, dtype=np.float32),
            "spring_damping": np.asarray(self._spring_kd, dtype=np.float32),
            "tri_indices": np.asarray(self._tri_idx, dtype=np.int32).reshape(-1, 3),
            "tri_poses": np.asarray(self._tri_pose, dtype=np.float32).reshape(-1, 2, 2),
            "tri_materials": np.asarray(self._tri_mat, dtype=np.float32).reshape(-1, 5),
            "tri_areas": np.asarray(self._tri_area, dtype=np.float32),
            "hex_indices": np.zeros((0, 8), np.int32),
            "hex_materials": np.zeros((0, 3), np.float32),
            "hex_inv_J0": np.zeros((0, 8, 3, 3), np.float32),
            "hex_det_J0_w": np.zeros((0, 8), np.float32),
        }

    def finalize(self, device="cuda", requires_grad: bool = False, particle_grid: bool = False) -> Model:
        """``particle_grid=True`` gives the model a :class:`HashGrid`, so that
        SolverFEM evaluates particle-particle contact."""
        m = Model(device, self._host_arrays(), self.gravity)
        if particle_grid:
            m.particle_grid = HashGrid()
        return m


class FEMBuilder(ModelBuilder):
    """Adds Hex8 elements with FP64 reference quadrature (reference builder.py:21-135)."""

    def __init__(self, *args, **kwargs):
        super().__init__(*args, **kwargs)
        self.fem_hex_indices = []
        self.fem_hex_materials = []

    def add_soft_mesh(self, pos, rot, scale, vel, vertices, indices, density, k_mu, k_lambda, k_damp, tri_ke=None,
                      tri_ka=None, tri_kd=None, tri_drag=None, tri_lift=None, add_surface_mesh_edges=True):
        """Tet4 mesh with lumped masses, open-face triangles and a unique-edge spring network.

        Particles at ``rot * (scale * v) + pos`` (``rot`` a quaternion x, y, z, w)
        with velocity ``vel`` and mass ``density * volume / 4`` from every
        incident tet; every face that belongs to one tet becomes a triangle
        membrane (Newton's ``add_soft_mesh``).  Each unique edge of a tet with
        positive rest volume receives one spring of stiffness ``k_mu / 2`` and
        damping ``k_damp / 2`` (reference builder.py:51-79).
        ``add_surface_mesh_edges`` is accepted for Newton compatibility: the
        bending edges it adds there are no force term of this solver (nor of
        the reference's), so none are stored.
        """
        verts = np.asarray(vertices, dtype=np.float64).reshape(-1, 3) * float(scale)
        x, y, z, w = (float(c) for c in np.asarray(rot, dtype=np.float64).reshape(4))
        R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                      [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                      [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
        start = self.particle_count
        pts = verts @ R.T + np.asarray(pos, dtype=np.float64).reshape(3)
        self.add_particles(pts, np.broadcast_to(np.asarray(vel, dtype=np.float64).reshape(3), pts.shape), 0.0)
        tets = np.asarray(indices, dtype=np.int64).reshape(-1, 4) + start
        first_tet = self.tet_count
        volume = self.add_tetrahedra(tets, [[k_mu, k_lambda, k_damp]])
        ok = volume > 0.0
        mass = np.asarray(self._m, dtype=np.float64)
        np.add.at(mass, tets[ok].ravel(), np.repeat(density * volume[ok] / 4.0, 4))
        self._m = mass.tolist()
        faces = {}
        for a, b, c, d in tets[ok].tolist():
            for face in ((a, c, b), (b, c, d), (a, b, d), (a, d, c)):
                key = tuple(sorted(face))
                if key in faces:
                    del faces[key]
                else:
                    faces[key] = face
        if faces:
            self.add_triangles(list(faces.values()), tri_ke, tri_ka, tri_kd, tri_drag, tri_lift)
        idx = np.concatenate(self._tet_idx)[first_tet:] if self._tet_idx else np.zeros((0, 4), np.int32)
        added, pairs = set(), []
        for a, b, c, d in idx.tolist():
            for left, right in ((a, b), (a, c), (a, d), (b, c), (b, d), (c, d)):
                key = (min(left, right), max(left, right))
                if key not in added:
                    added.add(key)
                    pairs.append(key)
        if pairs:
            self.add_springs(pairs, float(k_mu) * 0.5, float(k_damp) * 0.5, 0.0)

    def add_builder(self, builder, xform=None, label_prefix=None) -> None:
        """Copy another builder and remap its Hex8 indices (reference
        builder.py:36-49): a builder with Hex8 elements may be placed by a
        translation only, as Hex8 rest quadrature is taken from the final
        positions without rotating the elements' frames."""
        indices = getattr(builder, "fem_hex_indices", [])
        if indices and xform is not None:
            R, _ = _rigid(xform)
            if not np.allclose(R, np.eye(3), atol=1e-6):
                raise ValueError(
                    "Newton 1.7 particle replication supports translation only; rotate Hex8 rest vertices before adding them"
                )
        offset = self.particle_count
        super().add_builder(builder, xform=xform, label_prefix=label_prefix)
        for idx, mat in zip(indices, getattr(builder, "fem_hex_materials", [])):
            self.fem_hex_indices.append((np.asarray(idx) + offset).astype(np.int32))
            self.fem_hex_materials.append(np.array(mat))

    def add_soft_hex_mesh(self, hex_indices, k_mu=1.0e3, k_lambda=1.0e3, k_damp=0.0):
        indices = np.asarray(hex_indices, dtype=np.int32)
        if indices.ndim != 2 or indices.shape[1] != 8:
            raise ValueError("hex_indices must have shape (H, 8)")
        if indices.size and (indices.min() < 0 or indices.max() >= self.particle_count):
            raise ValueError("Hex8 indices must reference existing builder particles")
        materials = np.column_stack(
            [np.broadcast_to(np.asarray(v, dtype=np.float32), (len(indices),)) for v in (k_mu, k_lambda, k_damp)]
        )
        if not np.isfinite(materials).all() or np.any(materials < 0):
            raise ValueError("Hex8 materials must be finite and nonnegative")
        self.fem_hex_indices.append(indices)
        self.fem_hex_materials.append(materials)

    def finalize(self, device="cuda", requires_grad: bool = False, particle_grid: bool = False) -> Model:
        host = self._host_arrays()
        if self.fem_hex_indices:
            hexes = np.concatenate(self.fem_hex_indices).reshape(-1, 8)
            inv, det = hex_reference_data(np.asarray(self._q, dtype=np.float64), hexes)
            host["hex_indices"] = hexes.astype(np.int32)
            host["hex_materials"] = np.concatenate(self.fem_hex_materials).astype(np.float32).reshape(-1, 3)
            host["hex_inv_J0"] = inv.astype(np.float32)
            host["hex_det_J0_w"] = det.astype(np.float32)
        m = Model(device, host, self.gravity)
        if particle_grid:
            m.particle_grid = HashGrid()
        return m


def build_tet_model(scene, device="cuda", gravity=(0.0, 0.0, -9.81)) -> Model:
    """Model from a :class:`meshes.TetScene` (particles then tets, no springs)."""
    b = ModelBuilder(gravity=gravity)
    n = len(scene.points)
    b.add_particles(scene.points, np.zeros((n, 3)), scene.masses, radius=scene.radius)
    b.add_tetrahedra(scene.tets, scene.materials, activations=scene.activations)
    return b.finalize(device)


def build_hex_model(scene, device="cuda", gravity=(0.0, 0.0, -9.81)) -> Model:
    """Model from a :class:`meshes.HexScene`; masses ``rho * sum_g w_g N_a``."""
    b = FEMBuilder(gravity=gravity)
    pts = np.asarray(scene.points, dtype=np.float64)
    hexes = np.asarray(scene.hexes, dtype=np.int64)
    _, det = hex_reference_data(pts, hexes)
    gp = HEX_SIGNS / np.sqrt(3.0)
    shape = np.prod(1.0 + gp[:, None, :] * HEX_SIGNS[None, :, :], axis=2) / 8.0  # (8 points, 8 corners)
    masses = np.zeros(len(pts))
    np.add.at(masses, hexes.ravel(), (scene.density * det @ shape).ravel())
    b.add_particles(pts, np.zeros_like(pts), masses, radius=scene.radius)
    b.add_soft_hex_mesh(hexes, *scene.materials[0])
    return b.finalize(device)


def merge_tet_scenes(scenes):
    """Concatenate independent Tet4 scenes; returns (scene, particle offsets)."""
    from .meshes import TetScene

    offs = [0]
    pts, tets, mats, masses, fixed = [], [], [], [], []
    for s in scenes:
        base = offs[-1]
        pts.append(s.points)
        tets.append(s.tets + base)
        mats.append(s.materials)
        masses.append(s.masses)
        fixed.append(np.asarray(s.fixed, dtype=np.int64) + base)
        offs.append(base + len(s.points))
    merged = TetScene(
        np.concatenate(pts), np.concatenate(tets).astype(np.int32), np.concatenate(mats), np.concatenate(masses),
        fixed=np.concatenate(fixed).astype(np.int32), radius=scenes[0].radius,
    )
    return merged, np.asarray(offs, dtype=np.int64)
