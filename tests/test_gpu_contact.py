"""Particle-particle contact on cuda:0 (hash grid, through the C-ABI) vs the
FP64 oracle's all-pairs restatement (reference eval_particle_forces
kernels.py:1069-1114, eval_particle_particle_forces solver.py:1477-1510):

- forces of a random particle cloud (penalty, damping, friction, inactive
  particles, with and without a cohesion distance);
- the term's place in a step: a particle pile on the ground and two Tet4
  cubes pressed together, stepped against the oracle's step with the
  contact force given as a load;
- graph replays equal eager steps, repeated evaluations are bit-identical;
- models without a grid skip the term whatever self_contact_ke is.
"""

import numpy as np
import pytest
import torch

import oracle
from newton_solver_fem_amd import ModelBuilder, meshes
from test_gpu_parity import DEV, np_, rel, solver_for

pytestmark = pytest.mark.gpu

KE, KD, KF, MU = 400.0, 2.0, 50.0, 0.4


def cloud_model(n=3000, seed=0, cohesion=0.0, grid=True):
    rng = np.random.default_rng(seed)
    side = n ** (1.0 / 3.0) * 0.01
    q = rng.uniform(0.0, side, (n, 3)) + (0.0, 0.0, 0.02)
    v = rng.normal(0.0, 0.3, (n, 3))
    r = rng.uniform(0.003, 0.006, n)
    b = ModelBuilder(gravity=(0.0, 0.0, -9.81))
    for i in range(n):
        b.add_particle(q[i], v[i], mass=1.0e-3, radius=r[i], flags=0 if i % 11 == 0 else 1)
    m = b.finalize(DEV, particle_grid=grid)
    m.particle_kf, m.particle_mu, m.particle_cohesion = KF, MU, cohesion
    return m


def ref_contact(model, q, qd):
    h = model.host
    return oracle.particle_contact_forces(q, qd, h["particle_radius"], h["particle_flags"], model.particle_ke,
                                          model.particle_kd, model.particle_kf, model.particle_mu,
                                          model.particle_cohesion)


@pytest.mark.parametrize("cohesion", [0.0, 0.002])
def test_cloud_forces_match_oracle(cohesion):
    print(f"case contact cloud cohesion={cohesion}", flush=True)
    model = cloud_model(cohesion=cohesion)
    s = solver_for(model, dt=1e-3, self_contact_ke=KE, self_contact_kd=KD)
    st = model.state()
    f = s.eval_particle_particle_forces(model, None, st)
    ref = ref_contact(model, np_(st.particle_q), np_(st.particle_qd))
    touching = np.abs(ref).sum(1) > 0
    assert touching.sum() > 500  # the cloud has many contacts
    assert rel(np_(f), ref) < 2e-5
    flags = model.host["particle_flags"]
    assert np.all(np_(f)[flags == 0] == 0.0)
    g = s.eval_particle_particle_forces(model, None, st)
    assert torch.equal(f, g)  # fixed summation order: bit-identical


def test_pile_steps_match_oracle_and_graph_replays():
    print("case contact pile", flush=True)
    model = cloud_model(n=1500, seed=1)
    kw = dict(dt=1e-3, self_contact_ke=KE, self_contact_kd=KD, ground_plane=(0.0, 0.0, 1.0, 0.0), ground_ke=500.0,
              ground_kd=1.0, ground_kf=10.0, ground_mu=0.5, linear_solver_tolerance=1e-6, linear_solver_maxiter=50)
    s = solver_for(model, **kw)
    ref = oracle.OracleSolver(model.host, bits=64, dt=1e-3, gravity=model.gravity, ground_plane=(0, 0, 1, 0),
                              ground_ke=500.0, ground_kd=1.0, ground_kf=10.0, ground_mu=0.5, tolerance=1e-10,
                              maxiter=200)
    a, b = model.state(), model.state()
    q, qd = np_(a.particle_q), np_(a.particle_qd)
    for k in range(5):
        s.refresh_tangent(a.particle_q)
        ref.refresh_tangent(q)
        out = ref.step(q, qd, external_forces=ref_contact(model, q, qd))
        s.step(a, b, None, None, 1e-3)
        assert rel(np_(b.particle_qd), out["qd"]) < 1e-5, k
        assert rel(np_(b.particle_q), out["q"]) < 1e-6, k
        q, qd = np_(b.particle_q), np_(b.particle_qd)
        a, b = b, a
    # the same substeps recorded in a graph replay the eager ones bit for bit
    e0, e1 = model.state(), model.state()
    g0, g1 = model.state(), model.state()

    def body(x, y):
        s.refresh_tangent(x.particle_q)
        s.step(x, y, None, None, 1e-3)
        s.refresh_tangent(y.particle_q)
        s.step(y, x, None, None, 1e-3)

    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        body(g0, g1)
    torch.cuda.current_stream().wait_stream(side)
    graph = torch.cuda.CUDAGraph()
    g0.particle_q.copy_(model.particle_q)
    g0.particle_qd.copy_(model.particle_qd)
    with torch.cuda.graph(graph):
        body(g0, g1)
    for _ in range(3):
        g0.particle_q.copy_(model.particle_q)
        g0.particle_qd.copy_(model.particle_qd)
        s.dv.zero_()
        graph.replay()
        torch.cuda.synchronize()
        e0.particle_q.copy_(model.particle_q)
        e0.particle_qd.copy_(model.particle_qd)
        s.dv.zero_()
        body(e0, e1)
        torch.cuda.synchronize()
        assert torch.equal(g0.particle_q, e0.particle_q)
        assert torch.equal(g0.particle_qd, e0.particle_qd)


def two_cubes(gap=-0.002, grid=True):
    """Two 4x4x4-cell Tet4 cubes side by side along x, the second shifted so that
    their facing node layers are `gap` apart (negative: overlapping radii)."""
    n, h = 4, 0.01
    pts = meshes.lattice_points(n, n, n, h, origin=(0.0, 0.0, 0.05))
    tets = meshes.tet_lattice_indices(n, n, n)
    radius = 0.003  # intra-mesh neighbours (>= 0.01 apart) never touch
    b = ModelBuilder(gravity=(0.0, 0.0, 0.0))
    mu, lam = meshes.lame(1.0e5, 0.4)
    for k, x0 in enumerate((0.0, n * h + 2 * radius + gap)):
        off = b.particle_count
        p = pts + (x0, 0.0, 0.0)
        b.add_particles(p, np.tile((-0.2 if k else 0.2, 0.0, 0.0), (len(p), 1)),
                        meshes.tet_lumped_masses(p, tets, 1000.0), radius=radius)
        b.add_tetrahedra(tets + off, (mu, lam, 0.0))
    m = b.finalize(DEV, particle_grid=grid)
    m.particle_kf, m.particle_mu, m.particle_cohesion = KF, MU, 0.0
    return m


def test_fem_cubes_in_contact_step_matches_oracle():
    print("case contact cubes", flush=True)
    model = two_cubes()
    s = solver_for(model, dt=1e-3, self_contact_ke=2.0e3, self_contact_kd=5.0, linear_solver_tolerance=1e-6,
                   linear_solver_maxiter=200)
    ref = oracle.OracleSolver(model.host, bits=64, dt=1e-3, gravity=model.gravity, tolerance=1e-12, maxiter=400)
    a, b = model.state(), model.state()
    q, qd = np_(a.particle_q), np_(a.particle_qd)
    fc = ref_contact(model, q, qd)
    assert np.abs(fc).max() > 0  # the facing layers overlap
    for k in range(3):
        s.refresh_tangent(a.particle_q)
        ref.refresh_tangent(q)
        out = ref.step(q, qd, external_forces=ref_contact(model, q, qd))
        s.step(a, b, None, None, 1e-3)
        assert rel(np_(b.particle_qd), out["qd"]) < 1e-4, k
        q, qd = np_(b.particle_q), np_(b.particle_qd)
        a, b = b, a
    # The right-hand side holds h * contact force in the reference's term order.
    f = s.eval_particle_particle_forces(model, None, a)
    assert rel(np_(f), ref_contact(model, np_(a.particle_q), np_(a.particle_qd))) < 2e-5


def test_no_grid_skips_the_term():
    print("case contact no grid", flush=True)
    model = two_cubes(grid=False)
    s = solver_for(model, dt=1e-3, self_contact_ke=2.0e4, self_contact_kd=5.0)
    st = model.state()
    assert torch.count_nonzero(s.eval_particle_particle_forces(model, None, st)) == 0
    with_grid = two_cubes(grid=True)
    s2 = solver_for(with_grid, dt=1e-3, self_contact_ke=0.0)  # ke = 0 switches it off too
    assert torch.count_nonzero(s2.eval_particle_particle_forces(with_grid, None, with_grid.state())) == 0
