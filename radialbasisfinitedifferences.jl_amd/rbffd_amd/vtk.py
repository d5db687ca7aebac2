"""VTK output of node sets and fields, for visual checks of generated
operators and time-stepped fields (the reference writes its fields with
WriteVTK, src/processmesh.jl:29-31,108-109 and examples/adv_diff_test.jl:223-225).

Writes the legacy ASCII `.vtk` format (POLYDATA): the nodes as points, one
vertex cell per node unless triangles are given, and any number of named
scalar point fields."""
from __future__ import annotations

import numpy as np


def write_vtk(path, X, fields=None, triangles=None, title="rbffd fields"):
    """X: (N, 2) nodes; fields: {name: (N,) or (N, F) array}, a 2-D array
    writes F scalar fields name_0 .. name_{F-1}; triangles: (T, 3) node
    indices or None."""
    X = np.asarray(X, dtype=np.float64)
    N = len(X)
    fields = fields or {}
    with open(path, "w") as f:
        f.write("# vtk DataFile Version 3.0\n")
        f.write(title.replace("\n", " ")[:255] + "\n")
        f.write("ASCII\nDATASET POLYDATA\n")
        f.write(f"POINTS {N} double\n")
        np.savetxt(f, np.column_stack([X, np.zeros(N)]), fmt="%.17g")
        if triangles is None:
            f.write(f"VERTICES {N} {2 * N}\n")
            np.savetxt(f, np.column_stack([np.ones(N, np.int64), np.arange(N)]), fmt="%d")
        else:
            tri = np.asarray(triangles, dtype=np.int64)
            f.write(f"POLYGONS {len(tri)} {4 * len(tri)}\n")
            np.savetxt(f, np.column_stack([np.full(len(tri), 3), tri]), fmt="%d")
        if fields:
            f.write(f"POINT_DATA {N}\n")
            for name, v in fields.items():
                v = np.asarray(v, dtype=np.float64)
                cols = [v] if v.ndim == 1 else [v[:, j] for j in range(v.shape[1])]
                for j, c in enumerate(cols):
                    if len(c) != N:
                        raise ValueError(f"field {name} has {len(c)} values for {N} nodes")
                    label = name if v.ndim == 1 else f"{name}_{j}"
                    f.write(f"SCALARS {label} double 1\nLOOKUP_TABLE default\n")
                    np.savetxt(f, c, fmt="%.17g")


def read_vtk(path):
    """Points, cells and point fields of a file written by write_vtk."""
    with open(path) as f:
        toks = f.read().split("\n")
    i = 0
    pts, cells, fields = None, None, {}
    while i < len(toks):
        line = toks[i].split()
        if not line:
            i += 1
            continue
        if line[0] == "POINTS":
            n = int(line[1])
            pts = np.array([list(map(float, toks[i + 1 + k].split())) for k in range(n)])[:, :2]
            i += n + 1
        elif line[0] in ("VERTICES", "POLYGONS"):
            n = int(line[1])
            cells = np.array([list(map(int, toks[i + 1 + k].split()))[1:] for k in range(n)])
            i += n + 1
        elif line[0] == "SCALARS":
            name = line[1]
            n = len(pts)
            fields[name] = np.array([float(toks[i + 2 + k]) for k in range(n)])
            i += n + 2
        else:
            i += 1
    return pts, cells, fields
