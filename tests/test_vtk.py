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
