"""CPU restatement of the reserved-pool topology update (test infrastructure).

Checker only: used by tests, never by the product.  Pure-Python loops for
small pools.

- reserved_rows follows StreamingTetAssembly.__init__ (streaming.py:487-516):
  pool node i stands for original node i % count; row i holds i and every copy
  of every original neighbour of i % count in the valid tets.
- stage follows _stage_elements (streaming.py:34-131), _remove_elements
  (:134-158), _add_elements (:186-206), _compact_rows (:209-231) and the
  counter read-out of update() (:650-716): validation codes (the largest one
  wins, as the reference's atomic_max does), lineage (retained topology,
  reused blocks) and the update counters.  Threads of the reference race for a
  source's retained topology and cached blocks with compare-and-swap; here
  elements claim in ascending order, which yields the same counters (only the
  topology-claiming child of a source can also claim its blocks).

Pinned by the reference's own expectations in tests/test_reference_streaming.py
(see tests/test_oracle_streaming.py) and by brute force over the rows.
"""

from __future__ import annotations

import numpy as np

ERRORS = {
    1: "Element source is outside the previous element capacity",
    2: "Tetrahedral node is outside the reserved particle pool",
    3: "Cut introduces an edge outside the reserved original-node adjacency",
    4: "A tetrahedron must have four distinct original basis nodes",
}


def enabled(mat) -> bool:
    return bool(np.any(np.asarray(mat) != 0.0))


def reserved_rows(P: int, count: int, tets, mats) -> list[list[int]]:
    copies = P // count
    adj = [set() for _ in range(count)]
    for nodes, mat in zip(np.asarray(tets).reshape(-1, 4), np.asarray(mats).reshape(-1, 3)):
        if not enabled(mat):
            continue
        orig = [int(n) % count for n in nodes]
        for a in orig:
            adj[a].update(o for o in orig if o != a)
    return [sorted([i] + [nb + c * count for c in range(copies) for nb in adj[i % count]]) for i in range(P)]


def stage(rows, count, old, new, q_lin, q_now):
    """old: dict(nodes, poses, mats, active); new: dict(nodes, poses, mats,
    sources).  Returns (error_code, counters dict, new active flags)."""
    P = len(rows)
    T = len(new["nodes"])
    kept = [False] * T
    tkept = [False] * T
    retained = [-1] * T
    dirty = [False] * T
    active = [False] * T
    c = dict(dirty_elements=0, removed_elements=0, reused_elements=0, active_elements=0, touched_rows=0,
             topology_inserted_elements=0, topology_removed_elements=0)
    code = 0
    for e in range(T):
        if not enabled(new["mats"][e]):
            continue
        source = int(new["sources"][e])
        if source < -1 or source >= len(old["active"]):
            code = max(code, 1)
            continue
        nodes = [int(n) for n in new["nodes"][e]]
        bad = 0
        for a, n in enumerate(nodes):
            if n < 0 or n >= P:
                bad = 2
                break
            if any(n % count == nodes[b] % count for b in range(a)):
                bad = 4
                break
        if bad:
            code = max(code, bad)
            continue
        same = source >= 0 and bool(old["active"][source]) and nodes == [int(n) for n in old["nodes"][source]]
        if same:
            if not tkept[source]:
                tkept[source] = True
                retained[e] = source
        elif any(b not in rows[a] for a in nodes for b in nodes):
            code = max(code, 3)
            continue
        if retained[e] < 0:
            c["topology_inserted_elements"] += 1
            same = False
        if same:
            same = (all(np.array_equal(q_now[n], q_lin[n]) for n in nodes)
                    and np.array_equal(new["mats"][e], old["mats"][source])
                    and np.array_equal(new["poses"][e], old["poses"][source]))
        if same:
            same = not kept[source]
            kept[source] = True
        active[e] = True
        c["active_elements"] += 1
        if same:
            c["reused_elements"] += 1
        else:
            dirty[e] = True
            c["dirty_elements"] += 1
    if code:
        return code, None, None
    touched = set()
    for e in range(len(old["active"])):
        if not old["active"][e] or kept[e]:
            continue
        c["removed_elements"] += 1
        if not tkept[e]:
            c["topology_removed_elements"] += 1
            touched.update(int(n) for n in old["nodes"][e])
    for e in range(T):
        if dirty[e] and retained[e] < 0:
            touched.update(int(n) for n in new["nodes"][e])
    c["touched_rows"] = len(touched)
    c["full_numeric_refresh"] = c["dirty_elements"] > c["active_elements"] // 2
    return 0, c, active
