"""VTK output round trip: nodes, vertex or triangle cells, and fields."""
import numpy as np

from rbffd_amd.nodes import jittered_grid
from rbffd_amd.vtk import read_vtk, write_vtk


def test_vtk_round_trip(tmp_path):
    X = jittered_grid(7, 6, 0.25, seed=1)
    u = np.sin(X[:, 0]) * np.cos(X[:, 1])
    U = np.stack([u, 2 * u], 1)
    p = tmp_path / "f.vtk"
    write_vtk(p, X, {"u": u, "U": U})
    pts, cells, fields = read_vtk(p)
    assert np.array_equal(pts, X)
    assert np.array_equal(cells.ravel(), np.arange(len(X)))
    assert np.array_equal(fields["u"], u) and np.array_equal(fields["U_1"], 2 * u)
    from scipy.spatial import Delaunay

    tri = Delaunay(X).simplices
    write_vtk(p, X, {"u": u}, triangles=tri)
    _, cells, _ = read_vtk(p)
    assert np.array_equal(cells, tri)


# a file written by hand from the legacy-format specification: 3 points, a
# triangle and a vertex cell, two scalar fields (one without the optional
# component count)
LITERAL = """# vtk DataFile Version 3.0
hand-written check
ASCII
DATASET POLYDATA
POINTS 3 float
0 0 0  1 0 0
0 1 0
POLYGONS 1 4
3 0 1 2
POINT_DATA 3
SCALARS a double 1
LOOKUP_TABLE default
1.5 2.5 3.5
SCALARS b float
LOOKUP_TABLE default
0 1 2
"""


def test_spec_validator_on_a_literal_file():
    from vtk_spec import validate_legacy_polydata

    pts, cells, fields = validate_legacy_polydata(LITERAL)
    assert pts.shape == (3, 3) and cells == [[0, 1, 2]]
    assert np.array_equal(fields["a"], [1.5, 2.5, 3.5]) and np.array_equal(fields["b"], [0, 1, 2])


def test_writer_output_follows_the_spec(tmp_path):
    from vtk_spec import validate_legacy_polydata

    X = jittered_grid(6, 5, 0.25, seed=2)
    u = X[:, 0] * X[:, 1]
    p = tmp_path / "g.vtk"
    write_vtk(p, X, {"u": u, "U": np.stack([u, -u], 1)})
    pts, cells, fields = validate_legacy_polydata(p.read_text())
    assert np.array_equal(pts[:, :2], X) and cells == [[i] for i in range(len(X))]
    assert np.array_equal(fields["U_1"], -u)
