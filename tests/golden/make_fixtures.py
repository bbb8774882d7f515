"""Write the reference's ASCII mesh fixtures as JSON data files.

Reads ``examples/assets/{tet_grid,hex_grid,tet_cantilever}.usda`` from the
reference checkout (Apache-2.0, The Newton Developers) and stores only their
numeric arrays (points, element connectivity) in ``meshes.json`` next to this
script.  Run once in a container that has the reference mounted:

    python tests/golden/make_fixtures.py /root/reference
"""

import json
import re
import sys
from pathlib import Path


def _tuples(text, name):
    match = re.search(rf"{name}\s*=\s*\[(.*?)\]", text, re.S)
    body = match.group(1)
    return [[float(v) for v in t.split(",")] for t in re.findall(r"\(([^()]*)\)", body)]


def _ints(text, name):
    match = re.search(rf"{name}\s*=\s*\[(.*?)\]", text, re.S)
    return [int(v) for v in match.group(1).replace("(", "").replace(")", "").split(",") if v.strip()]


def main(root):
    assets = Path(root) / "src/newton_solver_fem/examples/assets"
    out = {}
    for name in ("tet_grid", "tet_cantilever"):
        text = (assets / f"{name}.usda").read_text()
        out[name] = {
            "points": _tuples(text, r"point3f\[\] points"),
            "tets": [[int(v) for v in t] for t in _tuples(text, r"int4\[\] tetVertexIndices")],
        }
    text = (assets / "hex_grid.usda").read_text()
    flat = _ints(text, "hex_vertex_indices")
    out["hex_grid"] = {
        "points": _tuples(text, r"point3f\[\] points"),
        "hexes": [flat[i : i + 8] for i in range(0, len(flat), 8)],
    }
    out["_source"] = "nvidia-beer/newton-solver-fem examples/assets (Apache-2.0, The Newton Developers)"
    (Path(__file__).parent / "meshes.json").write_text(json.dumps(out, indent=1))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
