"""The topology-update restatement in oracle/streaming_oracle.py against the
reference's own expectations (tests/test_reference_streaming.py:142-244) and
brute force.  CPU only."""

import numpy as np
import pytest

from helpers import POOL_CAPACITY, POOL_INITIAL, POOL_ORIGINAL, POOL_SPLIT, pool_arrays
import streaming_oracle as streaming

P = 3 * POOL_ORIGINAL


def _rest():
    from helpers import POOL_CORNERS

    return np.tile(POOL_CORNERS, (3, 1)).astype(np.float32)


def _state(records, sources=None):
    a = pool_arrays(records)
    src = np.arange(POOL_CAPACITY, dtype=np.int32) if sources is None else np.asarray(sources, dtype=np.int32)
    return {"nodes": a["tet_indices"], "poses": a["tet_poses"], "mats": a["tet_materials"], "sources": src}


def _sources(entries):
    values = np.full(POOL_CAPACITY, -1, dtype=np.int32)
    values[: len(entries)] = entries
    return values


def _published(records):
    s = _state(records)
    s["active"] = [streaming.enabled(m) for m in s["mats"]]
    return s


def _rows():
    a = pool_arrays(POOL_INITIAL)
    return streaming.reserved_rows(P, POOL_ORIGINAL, a["tet_indices"], a["tet_materials"])


def test_reserved_rows_are_every_copy_of_original_neighbours():
    a = pool_arrays(POOL_INITIAL)
    rows = _rows()
    # Brute force: (i, j) is reserved iff i == j or i % 5 and j % 5 share a valid tet.
    pairs = set()
    for t, m in zip(a["tet_indices"], a["tet_materials"]):
        if np.any(m != 0):
            pairs |= {(int(x) % 5, int(y) % 5) for x in t for y in t if x != y}
    for i in range(P):
        expect = [j for j in range(P) if j == i or (i % 5, j % 5) in pairs]
        assert rows[i] == expect
    # Original nodes 3 and 4 never share a tet: no copy of 4 in row 3.
    assert not any(j % 5 == 4 for j in rows[3])
    assert len(rows[0]) == 1 + 3 * 4


def test_split_shared_edge_removal_counters():
    rows = _rows()
    q = _rest()
    old = _published(POOL_INITIAL)
    code, c, _ = streaming.stage(rows, POOL_ORIGINAL, old, _state(POOL_SPLIT, _sources((1, 0, 0))), q, q)
    assert code == 0
    assert (c["dirty_elements"], c["reused_elements"], c["active_elements"]) == (2, 1, 3)
    assert c["topology_inserted_elements"] == 1 and c["topology_removed_elements"] == 0
    old = _published(POOL_SPLIT)
    code, c, _ = streaming.stage(rows, POOL_ORIGINAL, old, _state(POOL_SPLIT[:2]), q, q)
    assert code == 0 and c["dirty_elements"] == 0 and c["removed_elements"] == 1
    old = _published(POOL_SPLIT[:2])
    code, c, _ = streaming.stage(rows, POOL_ORIGINAL, old, _state(()), q, q)
    assert code == 0 and c["active_elements"] == 0 and c["removed_elements"] == 2


def test_moved_node_dirties_only_its_tet():
    rows = _rows()
    q = _rest()
    moved = q.copy()
    moved[4] += (0.15, -0.05, 0.09)
    code, c, _ = streaming.stage(rows, POOL_ORIGINAL, _published(POOL_INITIAL), _state(POOL_INITIAL), q, moved)
    assert code == 0
    assert (c["dirty_elements"], c["reused_elements"], c["touched_rows"]) == (1, 1, 0)
    assert c["topology_inserted_elements"] == 0 and c["topology_removed_elements"] == 0


@pytest.mark.parametrize("records,sources,code", [
    (((0, (0, 1, 3, 4), 1.0),), None, 3),          # nodes 3 and 4 were never adjacent
    (POOL_INITIAL, (POOL_CAPACITY, 1), 1),          # source outside the capacity
    (((0, (0, 6, 2, 3), 1.0),), None, 0),           # another copy of node 1: reserved
])
def test_validation_codes(records, sources, code):
    rows = _rows()
    q = _rest()
    new = _state(records, None if sources is None else _sources(sources))
    got, _, _ = streaming.stage(rows, POOL_ORIGINAL, _published(POOL_INITIAL), new, q, q)
    assert got == code


def _with_slot(state, slot, nodes):
    state["nodes"] = state["nodes"].copy()
    state["mats"] = state["mats"].copy()
    state["nodes"][slot] = nodes
    state["mats"][slot] = (1000.0, 1500.0, 0.0)
    return state


def test_repeated_original_node_and_pool_range():
    rows = _rows()
    q = _rest()
    old = _published(POOL_INITIAL)
    # 5 is a copy of 0 (coincident at rest, so the slot is written by hand).
    assert streaming.stage(rows, POOL_ORIGINAL, old, _with_slot(_state(POOL_INITIAL), 2, (0, 5, 2, 3)), q, q)[0] == 4
    assert streaming.stage(rows, POOL_ORIGINAL, old, _with_slot(_state(POOL_INITIAL), 2, (0, 1, 2, P)), q, q)[0] == 2
    # Several failures: the largest code is reported.
    new = _with_slot(_with_slot(_state(((0, (0, 1, 3, 4), 1.0),)), 2, (0, 5, 2, 3)), 3, (0, 1, 2, P))
    assert streaming.stage(rows, POOL_ORIGINAL, old, new, q, q)[0] == 4
