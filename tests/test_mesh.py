"""CGNS mesh import (src/processmesh.jl, src/extractelements.jl,
src/calculatenormal.jl, src/genghostnodes.jl) on the reference's mesh fixture
test/data/tominec_Y.cgns, read by this package's own HDF5 reader, and the
reference's mesh Poisson known answer (test/mesh_import_test.jl:158,
err < 0.001) with the oracle's operators."""
import os

import numpy as np
import pytest

from conftest import REF_DATA

MESH = os.path.join(REF_DATA, "tominec_Y.cgns")
MARKERS = ["dirichlet", "neumann"]


def test_hdf5_reader_reads_the_cgns_tree():
    from rbffd_amd.hdf5 import H5File

    f = H5File(MESH)
    assert set(f.children("Base/dom-1")) >= {"GridCoordinates", "TriElements", "dirichlet", "neumann", "ZoneBC"}
    x = f.read("Base/dom-1/GridCoordinates/CoordinateX/ data")
    y = f.read("Base/dom-1/GridCoordinates/CoordinateY/ data")
    assert x.dtype == np.float64 and x.shape == y.shape == (7773,)
    tri = f.read("Base/dom-1/TriElements/ElementConnectivity/ data")
    assert tri.shape == (45801,) and tri.min() == 1 and tri.max() == 7773   # 1-based, every vertex used
    assert list(f.read("Base/dom-1/TriElements/ElementRange/ data")) == [1, 15267]
    assert list(f.read("Base/dom-1/dirichlet/ElementRange/ data")) == [15268, 15396]
    assert list(f.read("Base/dom-1/neumann/ElementRange/ data")) == [15397, 15544]
    # element counts agree with the connectivity lengths
    assert f.read("Base/dom-1/dirichlet/ElementConnectivity/ data").shape == (2 * 129,)
    assert f.read("Base/dom-1/neumann/ElementConnectivity/ data").shape == (2 * 148,)
    assert bytes(f.read("Base/dom-1/ZoneType/ data")).decode() == "Unstructured"
    with pytest.raises(KeyError):
        f.read("Base/dom-1/nothing/ data")


def test_processmesh_nodes_normals_and_ghosts():
    from scipy.spatial import cKDTree

    from rbffd_amd.mesh import processmesh

    Y, V, yin, ybc, ybcg, cells, nrm, tng = processmesh(MESH, MARKERS)
    assert len(yin) == 15267 and [len(b) for b in ybc] == [129, 148]
    assert Y.shape == (15544 + 277, 2) and np.isfinite(Y).all()
    tri, lines = cells
    # nodes are triangle centroids and segment midpoints
    assert np.allclose(Y[yin], V[tri].mean(axis=1), rtol=0, atol=1e-15)
    assert np.allclose(np.concatenate([Y[b] for b in ybc]), V[lines].mean(axis=1), rtol=0, atol=1e-15)
    tree = cKDTree(Y[yin])
    h = np.mean(np.concatenate([tree.query(Y[b])[0] for b in ybc]))
    for b, g, n, t in zip(ybc, ybcg, nrm, tng):
        assert np.allclose(np.hypot(n[:, 0], n[:, 1]), 1.0) and np.allclose((n * t).sum(1), 0.0)
        # outward: the ghost node lies farther from the interior than its boundary node
        assert (tree.query(Y[g])[0] > tree.query(Y[b])[0]).mean() > 0.95
        assert np.allclose(np.hypot(*(Y[g] - Y[b]).T), h)


def _mesh_case(O, fitted_nodes):
    from rbffd_amd.mesh import processmesh

    X, _ = fitted_nodes
    Y_, _, yin, ybc, _, _, nrm, _ = processmesh(MESH, MARKERS)
    # test/mesh_import_test.jl:53-67: ghosts removed, the Y node nearest each X
    # node becomes that X node
    Y = np.concatenate([Y_[yin], Y_[ybc[0]], Y_[ybc[1]]])
    iy = O.knn(Y, X, 1)[0][:, 0]
    Y[iy] = X
    return X, Y, yin, ybc[1], ybc[0], nrm[1][:, 0], nrm[1][:, 1]


def test_mesh_poisson_known_answer(O, fitted_nodes):
    from test_oracle_weights import _csc, poisson_error

    X, Y, yin, yneu, ydir, xn, yn = _mesh_case(O, fitted_nodes)
    ops = O.generate_operator(X, Y, 3, 42, 5, ops=("E", "Dx", "Dy", "Dxx", "Dyy"))
    mats = [_csc(ops[k]) for k in ("E", "Dx", "Dy", "Dxx", "Dyy")]
    err = poisson_error(mats, X, Y, yin, yneu, ydir, xn, yn, O)
    print(f"mesh Poisson error with oracle operators: {err:.6f}")
    assert err < 0.001   # test/mesh_import_test.jl:158
