"""Topology updates (reserved duplicated-node pools and in-place re-sorting)
and component scaling on cuda:0, through the C-ABI, against a fresh
reconstruction of the edited model, the FP64 oracle and the lineage
restatement (reference tests/test_reference_streaming.py and
tests/test_reference_component_scaling.py).

Tolerances: matrices and inverse diagonals 3e-5 relative / 3e-6 absolute of
a fresh build (the reference's own), steps 2e-5 / 1e-4 as in its capture test;
patterns and update counters exact.
"""

import numpy as np
import pytest
import torch

import streaming_oracle
from helpers import (POOL_CAPACITY, POOL_INITIAL, POOL_ORIGINAL, POOL_SPLIT, cube, pool_arrays, pool_deformed,
                     pool_edit, pool_model)
from newton_solver_fem_amd import ModelBuilder, meshes
from newton_solver_fem_amd.model import Model
from test_gpu_parity import DEV, np_, oracle_for, rel, solver_for

pytestmark = pytest.mark.gpu

KW = dict(dt=1.0 / 120.0, linear_solver_maxiter=30)


def pool_solver(records=POOL_INITIAL, streaming=False, **kw):
    model = pool_model(records, DEV)
    args = dict(KW)
    args.update(kw)
    solver = solver_for(model, **args)
    if streaming:
        solver.prepare_tetrahedral_updates(POOL_ORIGINAL)
    return model, solver


def sources(entries):
    values = np.full(POOL_CAPACITY, -1, dtype=np.int32)
    values[: len(entries)] = entries
    return torch.tensor(values, device=DEV)


def storage(solver):
    return tuple(t.data_ptr() for t in solver._A_storage) + (solver._pc_inv_diag.data_ptr(), solver.dv.data_ptr())


def check_matrix(solver, positions, records):
    """A and M^-1 equal a fresh build of the edited model (reference
    _check_matrix) and the FP64 oracle's assembly."""
    model, fresh = pool_solver(records)
    fresh.refresh_tangent(positions)
    expected = fresh.A_bsr.dense()
    got = solver.A_bsr.dense()
    np.testing.assert_allclose(got, expected, rtol=3e-5, atol=3e-6)
    np.testing.assert_allclose(np_(solver._pc_inv_diag), np_(fresh._pc_inv_diag), rtol=3e-5, atol=3e-6)
    o = oracle_for(model, dt=KW["dt"])
    o.refresh_tangent(np_(positions))
    assert rel(got, o.dense()) < 1e-6


def published(solver_model, records):
    a = pool_arrays(records)
    return {"nodes": a["tet_indices"], "poses": a["tet_poses"], "mats": a["tet_materials"],
            "active": [streaming_oracle.enabled(m) for m in a["tet_materials"]]}


def expected_counters(rows, old_records, new_records, srcs, q_lin, q_now):
    a = pool_arrays(new_records)
    new = {"nodes": a["tet_indices"], "poses": a["tet_poses"], "mats": a["tet_materials"],
           "sources": np.arange(POOL_CAPACITY) if srcs is None else np.asarray(srcs)}
    code, c, _ = streaming_oracle.stage(rows, POOL_ORIGINAL, published(None, old_records), new, q_lin, q_now)
    assert code == 0
    return c


def test_split_shuffled_lineage_shared_edge_removal():
    print("case pool split/removal", flush=True)
    model, solver = pool_solver(streaming=True)
    P = model.particle_count
    a = pool_arrays(POOL_INITIAL)
    rows = streaming_oracle.reserved_rows(P, POOL_ORIGINAL, a["tet_indices"], a["tet_materials"])
    off = np.cumsum([0] + [len(r) for r in rows])
    np.testing.assert_array_equal(solver.A_bsr.offsets.cpu().numpy(), off)
    np.testing.assert_array_equal(solver.A_bsr.columns.cpu().numpy(), np.concatenate(rows))
    positions = pool_deformed(model)
    q = np_(positions).astype(np.float32)
    solver.refresh_tangent(positions)
    check_matrix(solver, positions, POOL_INITIAL)
    kept = storage(solver)

    src = sources((1, 0, 0))
    pool_edit(model, POOL_SPLIT)
    solver.refresh_topology(positions, element_sources=src)
    check_matrix(solver, positions, POOL_SPLIT)
    u = solver.last_update
    assert (u["dirty_elements"], u["reused_elements"], u["active_elements"]) == (2, 1, 3)
    want = expected_counters(rows, POOL_INITIAL, POOL_SPLIT, src.cpu().numpy(), q, q)
    assert u == want

    # The removed child shares three original nodes with the retained tet.
    retained = POOL_SPLIT[:2]
    pool_edit(model, retained)
    solver.refresh_topology(positions)
    check_matrix(solver, positions, retained)
    assert solver.last_update["dirty_elements"] == 0 and solver.last_update["removed_elements"] == 1
    assert solver.last_update == expected_counters(rows, POOL_SPLIT, retained, None, q, q)

    pool_edit(model, ())
    solver.refresh_topology(positions)
    check_matrix(solver, positions, ())
    assert solver.last_update["active_elements"] == 0
    assert storage(solver) == kept
    state, output = model.state(), model.state()
    solver.step(state, output, model.control(), None, solver.dt)
    np.testing.assert_array_equal(output.particle_q.cpu().numpy(), state.particle_q.cpu().numpy())
    np.testing.assert_array_equal(output.particle_qd.cpu().numpy(), 0.0)


def test_moved_unchanged_tet_refreshes_current_tangent():
    print("case pool moved node", flush=True)
    model, solver = pool_solver(streaming=True)
    positions = pool_deformed(model)
    solver.refresh_tangent(positions)
    moved = positions.clone()
    moved[4] += torch.tensor((0.15, -0.05, 0.09), device=DEV)
    solver.refresh_topology(moved)
    check_matrix(solver, moved, POOL_INITIAL)
    u = solver.last_update
    assert (u["dirty_elements"], u["reused_elements"], u["touched_rows"]) == (1, 1, 0)
    assert u["topology_inserted_elements"] == 0 and u["topology_removed_elements"] == 0


def test_reject_unreserved_edge_and_bad_source_without_mutating_assembly():
    print("case pool rejects", flush=True)
    model, solver = pool_solver(streaming=True)
    positions = pool_deformed(model)
    solver.refresh_tangent(positions)
    saved = [solver.A_bsr.values.clone(), solver._pc_inv_diag.clone()]
    # Original nodes 3 and 4 never shared an original tetrahedron.
    pool_edit(model, ((0, (0, 1, 3, 4), 1.0),))
    with pytest.raises(ValueError, match="outside the reserved original-node adjacency"):
        solver.refresh_topology(positions)
    for t, s in zip((solver.A_bsr.values, solver._pc_inv_diag), saved):
        assert torch.equal(t, s)
    pool_edit(model, POOL_INITIAL)
    with pytest.raises(ValueError, match="outside the previous element capacity"):
        solver.refresh_topology(positions, element_sources=sources((POOL_CAPACITY, 1)))
    assert torch.equal(solver.A_bsr.values, saved[0])
    solver.refresh_topology(positions)
    check_matrix(solver, positions, POOL_INITIAL)
    assert solver.last_update["dirty_elements"] == 0


def test_original_capture_replays_after_multiple_edits():
    print("case pool capture", flush=True)
    model, solver = pool_solver(streaming=True)
    state, output = model.state(), model.state()
    state.particle_q.copy_(pool_deformed(model))
    state.particle_qd.copy_(state.particle_q * torch.tensor((0.1, -0.05, 0.02), device=DEV))
    solver.refresh_tangent(state.particle_q)
    solver.step(state, output, None, None, solver.dt)
    kept = storage(solver)
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        solver.refresh_tangent(state.particle_q)
        solver.step(state, output, None, None, solver.dt)
    torch.cuda.current_stream().wait_stream(side)
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        solver.refresh_tangent(state.particle_q)
        solver.step(state, output, None, None, solver.dt)
    for records, srcs in ((POOL_SPLIT, (1, 0, 0)), (POOL_SPLIT[:2], (0, 1))):
        pool_edit(model, records)
        solver.refresh_topology(state.particle_q, element_sources=sources(srcs))
        g.replay()
        torch.cuda.synchronize()
        ref_model, ref = pool_solver(records)
        expected = ref_model.state()
        ref_in = ref_model.state()
        ref_in.particle_q.copy_(state.particle_q)
        ref_in.particle_qd.copy_(state.particle_qd)
        ref.refresh_tangent(ref_in.particle_q)
        ref.step(ref_in, expected, None, None, ref.dt)
        assert storage(solver) == kept
        np.testing.assert_allclose(np_(output.particle_q), np_(expected.particle_q), rtol=2e-5, atol=2e-6)
        np.testing.assert_allclose(np_(output.particle_qd), np_(expected.particle_qd), rtol=1e-4, atol=2e-5)
        o = oracle_for(ref_model, dt=KW["dt"], linear_solver_maxiter=30)
        o.refresh_tangent(np_(state.particle_q))
        r = o.step(np_(state.particle_q), np_(state.particle_qd))
        assert rel(np_(output.particle_qd), r["qd"]) < 2e-3


def lattice_pool(copies=3, seed=0):
    """A 2 x 2 x 2-cell Tet4 lattice (27 nodes, 40 tets) as a pool of `copies`
    coincident copies with room for 2 x 40 tets."""
    pts = np.asarray(meshes.lattice_points(2, 2, 2, 0.1), np.float64)
    tets = np.asarray(meshes.tet_lattice_indices(2, 2, 2), np.int64)
    n = len(pts)
    rest = np.tile(pts, (copies, 1))
    cap = 2 * len(tets)
    return rest, tets, n, cap


def lattice_arrays(rest, elems, cap):
    """Slots of (nodes, fraction) pairs: poses from the rest positions,
    materials and masses scaled by the fraction (unused slots zero)."""
    P = len(rest)
    idx = np.zeros((cap, 4), np.int32)
    poses = np.tile(np.eye(3), (cap, 1, 1))
    mats = np.zeros((cap, 3))
    masses = np.zeros(P)
    for slot, (nodes, frac) in enumerate(elems):
        dm = (rest[list(nodes[1:])] - rest[nodes[0]]).T
        idx[slot] = nodes
        poses[slot] = np.linalg.inv(dm)
        mats[slot] = frac * np.asarray((2000.0, 5000.0, 0.5))
        masses[list(nodes)] += frac * 500.0 * np.linalg.det(dm) / 24.0
    return {
        "tet_indices": idx, "tet_poses": poses.astype(np.float32), "tet_materials": mats.astype(np.float32),
        "particle_mass": np.where(masses > 0, masses, 1.0).astype(np.float32),
        "particle_inv_mass": np.divide(1.0, masses, out=np.zeros_like(masses), where=masses > 0).astype(np.float32),
        "particle_flags": (masses > 0).astype(np.int32),
    }


def lattice_model(rest, arrays):
    b = ModelBuilder(gravity=(0.0, 0.0, -9.81))
    for p in rest:
        b.add_particle(p, (0.0, 0.0, 0.0), mass=1.0, radius=0.0)
    host = b._host_arrays()
    host.update(arrays)
    host["tet_activations"] = np.zeros(len(arrays["tet_indices"]), np.float32)
    return Model(DEV, host, b.gravity)


def test_random_cuts_match_lineage_oracle_and_fresh_builds():
    """Random cuts of a 27-node lattice pool (three copies): corners moved to
    other copies, children splitting a parent's volume fraction, removals and
    shuffled slots with lineage; counters vs the restatement, A and M^-1 vs a
    fresh build and the FP64 oracle after every edit."""
    print("case pool random cuts", flush=True)
    rng = np.random.default_rng(11)
    rest, tets, n, cap = lattice_pool()
    elems = [(tuple(int(v) for v in t), 1.0) for t in tets]
    model = lattice_model(rest, lattice_arrays(rest, elems, cap))
    solver = solver_for(model, **KW)
    solver.prepare_tetrahedral_updates(n)
    P = len(rest)
    a0 = lattice_arrays(rest, elems, cap)
    rows = streaming_oracle.reserved_rows(P, n, a0["tet_indices"], a0["tet_materials"])
    q = (rest + rng.normal(scale=0.004, size=rest.shape)).astype(np.float32)
    positions = torch.tensor(q, device=DEV)
    solver.refresh_tangent(positions)
    old = {"nodes": a0["tet_indices"], "poses": a0["tet_poses"], "mats": a0["tet_materials"],
           "active": [streaming_oracle.enabled(m) for m in a0["tet_materials"]]}
    for round_ in range(4):
        new_elems, srcs = [], []
        for e, (nodes, frac) in enumerate(elems):
            r = rng.random()
            if r < 0.1:
                continue  # removed
            if r < 0.3 and len(elems) + len(new_elems) < cap:
                # split into two children on other copies of some corners
                for part in (0.5, 0.5):
                    moved = tuple(v % n + n * int(rng.integers(0, 3)) if rng.random() < 0.5 else v for v in nodes)
                    new_elems.append((moved, frac * part))
                    srcs.append(e)
                continue
            new_elems.append((nodes, frac))
            srcs.append(e if rng.random() < 0.9 else -1)
        order = rng.permutation(len(new_elems))
        new_elems = [new_elems[i] for i in order][:cap]
        srcs = [srcs[i] for i in order][:cap]
        arr = lattice_arrays(rest, new_elems, cap)
        for k, v in arr.items():
            getattr(model, k).copy_(torch.as_tensor(v).reshape(getattr(model, k).shape))
            model.host[k] = v
        src = np.full(cap, -1, np.int32)
        src[: len(srcs)] = srcs
        if round_ == 2:  # a moved node dirties its tets
            q[5] += (0.01, 0.0, -0.01)
            positions.copy_(torch.tensor(q))
        new = {"nodes": arr["tet_indices"], "poses": arr["tet_poses"], "mats": arr["tet_materials"], "sources": src}
        q_lin = np_(positions).astype(np.float32) if round_ != 2 else prev_q
        code, want, active = streaming_oracle.stage(rows, n, old, new, q_lin, q)
        assert code == 0
        solver.refresh_topology(positions, element_sources=torch.tensor(src, device=DEV))
        assert solver.last_update == want, (round_, solver.last_update, want)
        fresh_model = lattice_model(rest, arr)
        fresh = solver_for(fresh_model, **KW)
        fresh.refresh_tangent(positions)
        np.testing.assert_allclose(solver.A_bsr.dense(), fresh.A_bsr.dense(), rtol=3e-5, atol=3e-6)
        np.testing.assert_allclose(np_(solver._pc_inv_diag), np_(fresh._pc_inv_diag), rtol=3e-5, atol=3e-6)
        o = oracle_for(fresh_model, dt=KW["dt"])
        o.refresh_tangent(q)
        assert rel(solver.A_bsr.dense(), o.dense()) < 1e-6
        old = {"nodes": arr["tet_indices"], "poses": arr["tet_poses"], "mats": arr["tet_materials"], "active": active}
        elems = new_elems
        prev_q = q.copy()


def test_refresh_topology_without_pool_resorts_in_place():
    print("case re-sort in place", flush=True)
    model, solver = pool_solver()
    positions = pool_deformed(model)
    solver.refresh_tangent(positions)
    kept = storage(solver)
    pool_edit(model, POOL_SPLIT)
    solver.refresh_topology(positions)
    ref_model, ref = pool_solver(POOL_SPLIT)
    ref.refresh_tangent(positions)
    assert storage(solver) == kept
    np.testing.assert_array_equal(solver.A_bsr.offsets.cpu().numpy(), ref.A_bsr.offsets.cpu().numpy())
    np.testing.assert_array_equal(solver.A_bsr.columns.cpu().numpy(), ref.A_bsr.columns.cpu().numpy())
    np.testing.assert_allclose(solver.A_bsr.dense(), ref.A_bsr.dense(), rtol=3e-5, atol=3e-6)
    state, out, ref_out = model.state(), model.state(), ref_model.state()
    state.particle_q.copy_(positions)
    solver.step(state, out, None, None, solver.dt)
    ref_in = ref_model.state()
    ref_in.particle_q.copy_(positions)
    ref.step(ref_in, ref_out, None, None, ref.dt)
    np.testing.assert_allclose(np_(out.particle_qd), np_(ref_out.particle_qd), rtol=1e-4, atol=2e-5)


def _capture_substep(solver, state, output):
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        solver.refresh_tangent(state.particle_q)
        solver.step(state, output, None, None, solver.dt)
    torch.cuda.current_stream().wait_stream(side)
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        solver.refresh_tangent(state.particle_q)
        solver.step(state, output, None, None, solver.dt)
    return g


def _fresh_step(records, state):
    ref_model, ref = pool_solver(records)
    ref_in, expected = ref_model.state(), ref_model.state()
    ref_in.particle_q.copy_(state.particle_q)
    ref_in.particle_qd.copy_(state.particle_qd)
    ref.refresh_tangent(ref_in.particle_q)
    ref.step(ref_in, expected, None, None, ref.dt)
    return expected


def test_no_pool_edit_in_place_replays_captured_graph():
    """Without a pool an edit that fits the storage sized at construction is
    written into the same arrays and launches: a graph captured before the
    edit replays the edited topology (bit for bit what the solver now steps
    eagerly, and a fresh build's step to the solve tolerance) and the
    library allocates nothing."""
    print("case no-pool in-place capture", flush=True)
    model, solver = pool_solver()
    state, output = model.state(), model.state()
    state.particle_q.copy_(pool_deformed(model))
    state.particle_qd.copy_(state.particle_q * torch.tensor((0.1, -0.05, 0.02), device=DEV))
    g = _capture_substep(solver, state, output)
    kept, mem = storage(solver), solver.memory_stats()
    for records in (POOL_SPLIT, POOL_SPLIT[:2], POOL_INITIAL):
        pool_edit(model, records)
        solver.refresh_topology(state.particle_q)
        assert not solver.topology_reallocated
        assert solver.memory_stats() == mem and storage(solver) == kept
        g.replay()
        torch.cuda.synchronize()
        replayed_q, replayed_qd = np_(output.particle_q), np_(output.particle_qd)
        solver.dv.zero_()  # the replay started from the warm start the edit cleared
        solver.refresh_tangent(state.particle_q)
        solver.step(state, output, None, None, solver.dt)
        np.testing.assert_array_equal(replayed_q, np_(output.particle_q))
        np.testing.assert_array_equal(replayed_qd, np_(output.particle_qd))
        expected = _fresh_step(records, state)
        np.testing.assert_allclose(replayed_q, np_(expected.particle_q), rtol=2e-5, atol=2e-6)
        np.testing.assert_allclose(replayed_qd, np_(expected.particle_qd), rtol=1e-4, atol=2e-5)


def test_no_pool_edit_that_does_not_fit_reallocates_and_says_so():
    """The streamed CG plan is not re-planned in place: the edit gets new
    arrays, topology_reallocated is set, and the solver steps the edited
    topology (a fresh build's step to the solve tolerance)."""
    print("case no-pool reallocating edit", flush=True)
    model, solver = pool_solver()
    solver.set_cg_variant("streamed")
    state, output = model.state(), model.state()
    state.particle_q.copy_(pool_deformed(model))
    mem = solver.memory_stats()
    pool_edit(model, POOL_SPLIT)
    solver.refresh_topology(state.particle_q)
    assert solver.topology_reallocated
    assert solver.memory_stats()["allocations"] > mem["allocations"]
    solver.refresh_tangent(state.particle_q)
    solver.step(state, output, None, None, solver.dt)
    expected = _fresh_step(POOL_SPLIT, state)
    np.testing.assert_allclose(np_(output.particle_q), np_(expected.particle_q), rtol=2e-5, atol=2e-6)
    np.testing.assert_allclose(np_(output.particle_qd), np_(expected.particle_qd), rtol=1e-4, atol=2e-5)


def test_topology_argument_rules():
    print("case topology rules", flush=True)
    model, solver = pool_solver(streaming=True)
    positions = pool_deformed(model)
    with pytest.raises(RuntimeError, match="already prepared"):
        solver.prepare_tetrahedral_updates(POOL_ORIGINAL)
    with pytest.raises(ValueError, match="construction timestep"):
        solver.refresh_tangent(positions, KW["dt"] / 2)
    with pytest.raises(ValueError, match="currently assembled timestep"):
        solver.refresh_topology(positions, KW["dt"] / 2)
    with pytest.raises(ValueError, match="capacity-sized int32"):
        solver.refresh_topology(positions, element_sources=torch.zeros(2, dtype=torch.int32, device=DEV))
    with pytest.raises(ValueError, match="unscaled particle coordinates"):
        solver.prepare_step(model.state(), None, None, solver.dt)
    model.tet_materials = model.tet_materials.clone()
    with pytest.raises(ValueError, match="preserve model capacities"):
        solver.refresh_topology(positions)
    _, other = pool_solver()
    other.step(*(lambda m: (m.state(), m.state()))(other.model), None, None, other.dt)
    with pytest.raises(RuntimeError, match="before capturing or stepping"):
        other.prepare_tetrahedral_updates(POOL_ORIGINAL)
    _, fixed = pool_solver(fixed_particle_indices=(0,))
    with pytest.raises(ValueError, match="fixed_particle_indices"):
        fixed.prepare_tetrahedral_updates(POOL_ORIGINAL)
    with pytest.raises(ValueError, match="fixed_particle_indices"):
        fixed.prepare_component_scaling(torch.zeros(15, dtype=torch.int32, device=DEV), 1)
    _, bad = pool_solver()
    with pytest.raises(ValueError, match="whole copies"):
        bad.prepare_tetrahedral_updates(4)
    _, scaled = pool_solver()
    scaled.prepare_component_scaling(torch.zeros(15, dtype=torch.int32, device=DEV), 1)
    with pytest.raises(RuntimeError, match="before component scaling"):
        scaled.prepare_tetrahedral_updates(POOL_ORIGINAL)
    with pytest.raises(RuntimeError, match="already prepared"):
        scaled.prepare_component_scaling(torch.zeros(15, dtype=torch.int32, device=DEV), 1)
    _, fresh = pool_solver()
    with pytest.raises(ValueError, match="positive integer"):
        fresh.prepare_component_scaling(torch.zeros(15, dtype=torch.int32, device=DEV), 0)
    with pytest.raises(ValueError, match="particle-sized int32"):
        fresh.prepare_component_scaling(torch.zeros(14, dtype=torch.int32, device=DEV), 1)
    with pytest.raises(RuntimeError, match="Prepare component scaling before refreshing"):
        fresh.refresh_component_scaling()


# ---------------------------------------------------------------------------
# Component scaling
def test_equation_scaling_matches_its_definition():
    print("case scaling definition", flush=True)
    model, solver = pool_solver()
    labels = torch.tensor([0] * 5 + [-1] * 10, dtype=torch.int32, device=DEV)
    solver.prepare_component_scaling(labels, 6)
    state, out = model.state(), model.state()
    state.particle_q.copy_(pool_deformed(model))
    solver.refresh_tangent(state.particle_q)
    solver.step(state, out, None, None, solver.dt)
    torch.cuda.synchronize()
    cs = solver._component_scaling
    mass = model.host["particle_mass"].astype(np.float64)
    factors = np.ones(15, np.float32)
    factors[:5] = np.float32(1.0 / mass[:5].sum())
    np.testing.assert_allclose(cs.scale.cpu().numpy(), factors, rtol=1e-7)
    f3 = np.repeat(factors.astype(np.float64), 3)
    physical = solver.A_bsr.dense()
    scaled = cs.matrix.dense()
    np.testing.assert_allclose(scaled, physical * f3[:, None], rtol=1e-7, atol=0)
    np.testing.assert_allclose(scaled, scaled.T, rtol=0, atol=2e-6 * np.abs(scaled).max())
    np.testing.assert_allclose(np_(cs.inverse), np_(solver._pc_inv_diag) / factors[:, None, None], rtol=1e-6)
    np.testing.assert_allclose(np_(cs.rhs), np_(state.particle_f) * factors[:, None], rtol=1e-6)
    np.testing.assert_array_equal(model.particle_mass.cpu().numpy(), model.host["particle_mass"])


def two_cubes(scale2, stiff2=1.0):
    """Two disjoint Tet4 cubes; the second one's masses are multiplied by
    scale2 and its materials by scale2 * stiff2 (stiff2 != 1 gives it another
    spectrum)."""
    m = cube("tet", device="cpu")
    h = m.host
    pts = h["particle_q"].astype(np.float64)
    tets = h["tet_indices"].astype(np.int64)
    n = len(pts)
    b = ModelBuilder(gravity=(0.0, 0.0, -9.81))
    masses = h["particle_mass"].astype(np.float64)
    b.add_particles(pts, np.zeros_like(pts), masses, radius=0.0)
    b.add_particles(pts + (0.3, 0.0, 0.0), np.zeros_like(pts), masses * scale2, radius=0.0)
    b.add_tetrahedra(tets, (500.0, 1000.0, 0.0))
    b.add_tetrahedra(tets + n, (500.0 * scale2 * stiff2, 1000.0 * scale2 * stiff2, 0.0))
    return b.finalize(DEV), n


def test_fixed_budget_pcg_resolves_tiny_disconnected_component():
    """A component 1e-20 times lighter (and with another stiffness-to-mass
    ratio): with equation scaling an 80-iteration budget solves both
    components; without it the light one is left unresolved (reference test_fixed_iteration_pcg_resolves_tiny_disconnected_block)."""
    print("case scaling tiny component", flush=True)
    model, n = two_cubes(1e-20, 40.0)
    qd = np.random.default_rng(3).normal(scale=0.05, size=(2 * n, 3)).astype(np.float32)
    q = model.host["particle_q"].copy()
    q[:, 0] += 0.05 * q[:, 2]
    # Exact dv of each component: the oracle at normal scale (the light cube's
    # equations are the normal cube's times 1e-20).
    normal, _ = two_cubes(1.0, 40.0)
    o = oracle_for(normal, dt=0.01, linear_solver_tolerance=1e-12, linear_solver_maxiter=2000)
    o.refresh_tangent(q)
    exact = o.step(q, qd)["qd"] - qd.astype(np.float64)
    errs = {}
    for scaled in (True, False):
        s = solver_for(model, dt=0.01, linear_solver_maxiter=80, linear_solver_tolerance=1e-6)
        if scaled:
            s.prepare_component_scaling(torch.tensor([0] * n + [1] * n, dtype=torch.int32, device=DEV), 2)
        state, out = model.state(), model.state()
        state.particle_q.copy_(torch.tensor(q))
        state.particle_qd.copy_(torch.tensor(qd))
        s.refresh_tangent(state.particle_q)
        s.step(state, out, None, None, s.dt)
        dv = np_(out.particle_qd) - qd
        errs[scaled] = (rel(dv[:n], exact[:n]), rel(dv[n:], exact[n:]))
    print("relative dv errors (scaled, unscaled):", errs, flush=True)
    assert errs[True][0] < 1e-2 and errs[True][1] < 1e-2
    assert errs[False][1] > 10 * errs[True][1]


def test_mass_and_label_updates_reuse_buffers():
    print("case scaling refresh", flush=True)
    model, solver = pool_solver()
    labels = torch.tensor([0] * 5 + [-1] * 10, dtype=torch.int32, device=DEV)
    solver.prepare_component_scaling(labels, 3)
    cs = solver._component_scaling
    pointers = (cs.scale.data_ptr(), cs.total_mass.data_ptr(), cs._values.data_ptr(), cs.inverse.data_ptr())
    labels.copy_(torch.tensor([0] * 4 + [-1, 2, 2, 2] + [5] * 7, dtype=torch.int32))
    model.particle_mass.copy_(torch.tensor([0.5] * 4 + [1.0, 2e-22, 3e-22, 5e-22] + [1.0] * 7))
    solver.refresh_component_scaling()
    np.testing.assert_allclose(cs.scale.cpu().numpy(), [0.5] * 4 + [1.0] + [1e21] * 3 + [1.0] * 7, rtol=1e-7)
    assert pointers == (cs.scale.data_ptr(), cs.total_mass.data_ptr(), cs._values.data_ptr(), cs.inverse.data_ptr())


def test_capture_uses_refreshed_component_factors():
    """A step recorded once replays with factors refreshed between replays
    and equals an eager solver refreshed the same way (same kernels)."""
    print("case scaling capture", flush=True)
    model, n = two_cubes(1e-3)
    labels = torch.tensor([0] * n + [1] * n, dtype=torch.int32, device=DEV)
    qd = torch.tensor(np.random.default_rng(4).normal(scale=0.05, size=(2 * n, 3)), dtype=torch.float32, device=DEV)
    solvers = []
    for _ in range(2):
        s = solver_for(model, dt=0.01, linear_solver_maxiter=20, linear_solver_tolerance=1e-4)
        s.prepare_component_scaling(labels, 2)
        solvers.append(s)
    cap, eager = solvers
    state, out = model.state(), model.state()
    state.particle_qd.copy_(qd)
    cap.refresh_tangent(state.particle_q)
    eager.refresh_tangent(state.particle_q)
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        cap.step(state, out, None, None, cap.dt)
    torch.cuda.current_stream().wait_stream(side)
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        cap.step(state, out, None, None, cap.dt)
    base = model.particle_mass.clone()
    for factor in (0.5, 1.0, 2.0):
        model.particle_mass.copy_(base * factor)
        for s in solvers:
            s.refresh_component_scaling()
            s.dv.zero_()
        g.replay()
        ref_in, ref_out = model.state(), model.state()
        ref_in.particle_qd.copy_(qd)
        eager.step(ref_in, ref_out, None, None, eager.dt)
        torch.cuda.synchronize()
        np.testing.assert_allclose(np_(cap._component_scaling.scale), np_(eager._component_scaling.scale), rtol=0)
        np.testing.assert_allclose(np_(out.particle_qd), np_(ref_out.particle_qd), rtol=1e-6, atol=1e-7)
    model.particle_mass.copy_(base)


@pytest.mark.parametrize("streaming", [False, True])
def test_topology_refresh_retains_scaled_connectivity_storage(streaming):
    print(f"case scaling topology streaming={streaming}", flush=True)
    model, solver = pool_solver(streaming=streaming)
    labels = torch.tensor([0] * 5 + [-1] * 10, dtype=torch.int32, device=DEV)
    solver.prepare_component_scaling(labels, 6)
    current, following = model.state(), model.state()
    deformed = pool_deformed(model)
    current.particle_q.copy_(deformed)
    solver.refresh_tangent(deformed)
    solver.step(current, following, None, None, solver.dt)
    cs = solver._component_scaling
    pointers = (cs.matrix.offsets.data_ptr(), cs.matrix.columns.data_ptr(), cs._values.data_ptr())
    pool_edit(model, POOL_SPLIT)
    labels.copy_(torch.tensor([0] * 5 + [1] * 4 + [-1] * 6, dtype=torch.int32))
    solver.refresh_component_scaling()
    solver.refresh_topology(deformed)
    solver.step(current, following, None, None, solver.dt)
    torch.cuda.synchronize()
    assert cs.matrix.offsets.data_ptr() == solver.A_bsr.offsets.data_ptr()
    assert cs.matrix.columns.data_ptr() == solver.A_bsr.columns.data_ptr()
    assert pointers == (cs.matrix.offsets.data_ptr(), cs.matrix.columns.data_ptr(), cs._values.data_ptr())
    f3 = np.repeat(cs.scale.cpu().numpy().astype(np.float64), 3)
    np.testing.assert_allclose(cs.matrix.dense(), solver.A_bsr.dense() * f3[:, None], rtol=1e-7, atol=0)
    assert np.all(np.isfinite(np_(following.particle_q)))


@pytest.mark.parametrize("method", ["refresh_tangent", "refresh_topology", "rebuild_element_blocks"])
def test_reassembly_uses_the_same_fixed_positions_as_force_evaluation(method):
    """Reference tests/test_fixed_particles.py:122-140: every reassembly entry
    restores the anchors of fixed particles in the positions it is handed."""
    print(f"case fixed reassembly {method}", flush=True)
    from helpers import tet_pair

    model = tet_pair(DEV)
    kw = dict(dt=0.002, fixed_particle_indices=(0, 1), linear_solver_maxiter=64)
    expected, actual = solver_for(model, **kw), solver_for(model, **kw)
    reference, changed = model.state(), model.state()
    positions = model.particle_q.cpu().numpy().copy()
    positions[2] += (0.002, 0.001, -0.001)
    reference.particle_q.copy_(torch.tensor(positions))
    positions[:2] += (0.03, -0.02, 0.04)
    changed.particle_q.copy_(torch.tensor(positions))
    getattr(expected, method)(reference.particle_q, 0.002)
    getattr(actual, method)(changed.particle_q, 0.002)
    np.testing.assert_array_equal(changed.particle_q.cpu().numpy(), reference.particle_q.cpu().numpy())
    np.testing.assert_array_equal(actual.A_bsr.dense(), expected.A_bsr.dense())
    expected_out, actual_out = model.state(), model.state()
    expected.step(reference, expected_out, None, None, 0.002)
    actual.step(changed, actual_out, None, None, 0.002)
    np.testing.assert_allclose(np_(actual_out.particle_qd), np_(expected_out.particle_qd), atol=1e-7)


def test_scaled_pool_capture_replays_after_edit():
    """A step recorded with a reserved pool and component scaling (scaled
    copies gathered into the compact Krylov pattern) replays after an edit,
    refreshed labels and factors like an eager solver built for the edit."""
    print("case scaled pool capture", flush=True)
    model, solver = pool_solver(streaming=True)
    labels = torch.tensor([0] * 5 + [-1] * 10, dtype=torch.int32, device=DEV)
    solver.prepare_component_scaling(labels, 6)
    state, output = model.state(), model.state()
    state.particle_q.copy_(pool_deformed(model))
    state.particle_qd.copy_(state.particle_q * torch.tensor((0.1, -0.05, 0.02), device=DEV))
    solver.refresh_tangent(state.particle_q)
    solver.step(state, output, None, None, solver.dt)
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        solver.refresh_tangent(state.particle_q)
        solver.step(state, output, None, None, solver.dt)
    torch.cuda.current_stream().wait_stream(side)
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        solver.refresh_tangent(state.particle_q)
        solver.step(state, output, None, None, solver.dt)
    split_labels = [0] * 5 + [1] * 4 + [-1] * 6
    pool_edit(model, POOL_SPLIT)
    labels.copy_(torch.tensor(split_labels, dtype=torch.int32))
    solver.refresh_component_scaling()
    solver.refresh_topology(state.particle_q, element_sources=sources((1, 0, 0)))
    g.replay()
    torch.cuda.synchronize()
    ref_model, ref = pool_solver(POOL_SPLIT, streaming=True)
    ref.prepare_component_scaling(torch.tensor(split_labels, dtype=torch.int32, device=DEV), 6)
    ref_in, expected = ref_model.state(), ref_model.state()
    ref_in.particle_q.copy_(state.particle_q)
    ref_in.particle_qd.copy_(state.particle_qd)
    ref.refresh_tangent(ref_in.particle_q)
    ref.step(ref_in, expected, None, None, ref.dt)
    np.testing.assert_allclose(np_(output.particle_q), np_(expected.particle_q), rtol=2e-5, atol=2e-6)
    np.testing.assert_allclose(np_(output.particle_qd), np_(expected.particle_qd), rtol=1e-4, atol=2e-5)
    np.testing.assert_allclose(np_(solver._component_scaling.scale), np_(ref._component_scaling.scale), rtol=1e-7)


def test_unchanged_topology_refresh_keeps_every_array():
    """A refresh_topology without a pool and without an edit re-derives the
    same pattern and keeps the matrix and its storage (the library allocates
    nothing in that case, so per-frame refreshes do not grow memory)."""
    print("case unchanged refresh", flush=True)
    model, solver = pool_solver()
    positions = pool_deformed(model)
    solver.refresh_tangent(positions)
    before = solver.A_bsr.dense()
    pointers = storage(solver)
    mem = solver.memory_stats()
    for _ in range(5):
        solver.refresh_topology(positions)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(solver.A_bsr.dense(), before)
    assert storage(solver) == pointers
    assert solver.memory_stats() == mem  # the library allocated nothing either
