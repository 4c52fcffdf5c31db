"""Write the field names of the reference's solver types (build container only: it parses the
reference's sources under ``/root/reference/src`` as text, nothing is imported).

``ref_interface_fields.json``: for ``Ops3D``, ``State3D`` (ns3d/core.py), ``Geometry3D``
(ns3d/geometry.py), ``SolverConfig3D`` (ns3d/config.py), ``Ops``, ``State`` (ns2d/core.py) and
``Geometry`` (ns2d/geometry.py), the annotated fields in declaration order and the names of the
methods / properties. The seam tests build objects carrying exactly these attributes.

Usage: ``python tests/golden/make_interface_fields.py``.
"""
from __future__ import annotations

import ast
import json
from pathlib import Path

SRC = Path("/root/reference/src/aerolume/solvers")
OUT = Path(__file__).resolve().parent / "ref_interface_fields.json"
WANT = {"ns3d/core.py": ("Ops3D", "State3D"), "ns3d/geometry.py": ("Geometry3D",),
        "ns3d/config.py": ("SolverConfig3D",), "ns2d/core.py": ("Ops", "State"), "ns2d/geometry.py": ("Geometry",)}


def main():
    out = {}
    for rel, names in WANT.items():
        tree = ast.parse((SRC / rel).read_text())
        for node in tree.body:
            if isinstance(node, ast.ClassDef) and node.name in names:
                fields = [s.target.id for s in node.body if isinstance(s, ast.AnnAssign)]
                methods = [f.name for f in node.body if isinstance(f, ast.FunctionDef)]
                out[node.name] = {"fields": fields, "methods": methods}
    OUT.write_text(json.dumps(out, indent=1) + "\n")
    print("wrote", OUT, sorted(out))


if __name__ == "__main__":
    main()
