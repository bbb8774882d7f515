"""ModelBuilder / FEMBuilder replication (Newton's add_builder, reference
builder.py:36-49), checked on the host (no GPU):

- the reference's tests/test_reference_builder.py:65-84: replicated Hex8
  indices are remapped, a rotated placement of a Hex8 builder is refused, a
  translated one keeps the Hex8 quadrature (det J0 w = 3) and places the
  nodes at the transformed positions;
- a Tet4 + spring + triangle builder placed by a rotation and a translation:
  indices offset, positions and velocities transformed, element rest data
  (poses, rest lengths, areas, materials) copied, and the copy's deformation
  gradients are the originals rotated (F' = R F), so its elastic energy is
  unchanged.
"""

import numpy as np
import pytest

from newton_solver_fem_amd import FEMBuilder, ModelBuilder

NODES = np.asarray([(0, 0, 0), (2, 0, 0), (2, 3, 0), (0, 3, 0), (0, 0, 4), (2, 0, 4), (2, 3, 4), (0, 3, 4)], float)


def quat(axis, angle):
    axis = np.asarray(axis, float) / np.linalg.norm(axis)
    return (*(axis * np.sin(angle / 2)), np.cos(angle / 2))


def hex_builder():
    b = FEMBuilder()
    b.add_particles(NODES, np.zeros_like(NODES), [3.0] * 8)
    b.add_soft_hex_mesh(np.arange(8).reshape(1, 8), k_mu=10.0, k_lambda=20.0, k_damp=0.0)
    return b


def test_hex_replication_remaps_indices_and_rejects_rotation():
    local = hex_builder()
    combined = FEMBuilder()
    combined.add_builder(local)
    with pytest.raises(ValueError, match="translation only"):
        combined.add_builder(local, xform=(10.0, 20.0, 30.0, *quat((0, 0, 1), 0.7)))
    combined.add_builder(local, xform=(10.0, 20.0, 30.0, 0.0, 0.0, 0.0, 1.0))
    model = combined.finalize(device="cpu")
    np.testing.assert_array_equal(model.host["hex_indices"], np.arange(16).reshape(2, 8))
    np.testing.assert_allclose(model.host["hex_det_J0_w"], 3.0, rtol=2e-6)
    np.testing.assert_allclose(model.host["particle_q"][8:], NODES + (10.0, 20.0, 30.0), atol=2e-6)


def tet_builder():
    b = ModelBuilder(gravity=(0.0, 0.0, -9.81))
    pts = np.array([(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 1, 1)], float) * 0.1
    b.add_particles(pts, np.tile((0.1, -0.2, 0.3), (5, 1)), [1.0, 2.0, 3.0, 4.0, 5.0], radius=0.01)
    b.add_tetrahedra([[0, 1, 2, 3], [1, 2, 3, 4]], [[100.0, 200.0, 1.0]])
    b.add_springs([(0, 4), (1, 3)], 50.0, 0.5)
    b.add_triangles([(0, 1, 2), (1, 2, 4)], tri_ke=10.0, tri_ka=20.0)
    return b


def test_tet_builder_replicated_by_a_rigid_transform():
    local = tet_builder()
    R_q = quat((1, 2, 3), 1.1)
    shift = np.array([0.5, -0.25, 2.0])
    combined = ModelBuilder(gravity=(0.0, 0.0, -9.81))
    combined.add_builder(local)
    combined.add_builder(local, xform=(*shift, *R_q))
    a = local._host_arrays()
    h = combined._host_arrays()
    n = len(a["particle_q"])
    from newton_solver_fem_amd.model import _rigid

    R, _ = _rigid((0, 0, 0, *R_q))
    np.testing.assert_allclose(R @ R.T, np.eye(3), atol=1e-12)
    np.testing.assert_allclose(h["particle_q"][n:], a["particle_q"] @ R.T + shift, atol=1e-6)
    np.testing.assert_allclose(h["particle_qd"][n:], a["particle_qd"] @ R.T, atol=1e-6)
    np.testing.assert_array_equal(h["particle_mass"][n:], a["particle_mass"])
    for key, width in (("tet_indices", 4), ("spring_indices", 2), ("tri_indices", 3)):
        np.testing.assert_array_equal(h[key][len(a[key]):], a[key] + n)
    for key in ("tet_poses", "tet_materials", "spring_rest_length", "spring_stiffness", "tri_poses", "tri_areas",
                "tri_materials"):
        np.testing.assert_array_equal(h[key][len(a[key]):], a[key])
    # deformation gradients of the copy at rest: F' = Ds' Dm^-1 = R (Ds Dm^-1) = R F
    q2 = h["particle_q"].astype(np.float64)
    for t, (idx, pose) in enumerate(zip(h["tet_indices"][len(a["tet_indices"]):], h["tet_poses"][2:])):
        ds = (q2[idx[1:]] - q2[idx[0]]).T
        F = ds @ pose.astype(np.float64)
        np.testing.assert_allclose(F, R, atol=1e-5)  # the original is at rest: F = I
    # rest lengths still match the transformed positions
    sp = h["spring_indices"][2:]
    np.testing.assert_allclose(np.linalg.norm(q2[sp[:, 0]] - q2[sp[:, 1]], axis=1), h["spring_rest_length"][2:],
                               rtol=1e-6)
