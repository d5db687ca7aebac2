"""VTK output of a GPU result, checked against the legacy VTK file format
(the "Simple Legacy Formats" of the VTK file-format specification) by a
token-level validator written here from that specification, independent of
rbffd_amd.vtk.read_vtk: header line, title, ASCII, DATASET POLYDATA, POINTS
n type with 3 n coordinates, a cell section whose size equals the sum of
(1 + points) over its cells with indices in range, and POINT_DATA n with
SCALARS name type 1 / LOOKUP_TABLE default / n values per field.  The field
is the GPU Laplacian of u = sin(pi x) sin(pi y) applied on the GPU
(the reference writes its fields with WriteVTK, src/processmesh.jl:29-31)."""
import numpy as np
import pytest

from rbffd_amd.nodes import jittered_grid

pytestmark = pytest.mark.gpu

from vtk_spec import validate_legacy_polydata


def test_vtk_of_gpu_field(ctx, tmp_path):
    import rbffd_amd
    from rbffd_amd.vtk import write_vtk

    X = jittered_grid(40, 41, 0.25, seed=5)
    n = len(X)
    print("VTK of a GPU Laplacian", flush=True)
    eng = rbffd_amd.Engine(ctx)
    Xd = ctx.to_device(X)
    col, vals, _ = eng.generate(Xd, Xd, 5, 13, 2, ["Lap"])
    u = np.sin(np.pi * X[:, 0]) * np.sin(np.pi * X[:, 1])
    y = ctx.empty((n,), np.float64)
    eng.spmv(col, vals[0], ctx.to_device(u), y, uniform_n=13)
    lap = y.get()
    path = tmp_path / "lap.vtk"
    write_vtk(path, X, {"u": u, "lap_u": lap})
    pts, cells, fields = validate_legacy_polydata(path.read_text())
    assert np.array_equal(pts[:, :2], X) and not pts[:, 2].any()
    assert cells == [[i] for i in range(n)]
    assert np.array_equal(fields["u"], u) and np.array_equal(fields["lap_u"], lap)
    inner = (X[:, 0] > 0.1) & (X[:, 0] < 0.9) & (X[:, 1] > 0.1) & (X[:, 1] < 0.9)
    err = np.abs(lap[inner] + 2 * np.pi**2 * u[inner]).max() / (2 * np.pi**2)
    print(f"  interior relative error of the Laplacian {err:.2e}", flush=True)
    assert err < 0.05
    from scipy.spatial import Delaunay

    tri = Delaunay(X).simplices
    write_vtk(path, X, {"lap_u": lap}, triangles=tri)
    _, cells, _ = validate_legacy_polydata(path.read_text())
    assert np.array_equal(np.array(cells), tri)
