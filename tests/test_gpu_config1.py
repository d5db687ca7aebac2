"""The CPU-runnable configuration: 2D Poisson on [0,1]^2, 51 x 51 nodes (h = 0.02,
interior jittered +-0.25h, seed 1), r^5, degree 2, n = 13, manufactured solution
u = sin(pi x) sin(pi y) with Dirichlet rows.  GPU weights against the oracle,
and the L2 error of the sparse direct solve with GPU operators equal to the one
with oracle operators."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as sla

from rbffd_amd.nodes import config_nodes

pytestmark = pytest.mark.gpu


def _poisson_l2(X, Lap, E):
    x, y = X[:, 0], X[:, 1]
    u = np.sin(np.pi * x) * np.sin(np.pi * y)
    f = -2 * np.pi**2 * u
    edge = (x <= 0) | (x >= 1) | (y <= 0) | (y >= 1)
    A = sp.lil_matrix(Lap)
    A[np.where(edge)[0], :] = E[np.where(edge)[0], :]
    rhs = np.where(edge, u, f)
    sol = sla.spsolve(A.tocsc(), rhs)
    return float(np.linalg.norm(sol - u) / np.linalg.norm(u))


def test_config1_poisson(ctx, O):
    import rbffd_amd

    X, prm = config_nodes("poisson2500")
    p, n, deg = prm["p"], prm["n"], prm["polydeg"]
    print("config 1: 51 x 51 Poisson, r^5, degree 2, n = 13", flush=True)
    E, Lap = rbffd_amd.operators(X, X, p, n, deg, ["E", "Lap"], ctx=ctx)
    ref = O.generate_operator(X, X, p, n, deg, ops=("E", "Lap"), quad=True)
    for name, A in (("E", E), ("Lap", Lap)):
        cp, rv, nz, shape = ref[name]
        B = sp.csc_matrix((nz, rv, cp), shape=shape)
        assert A.shape == B.shape
        assert np.array_equal(A.indptr, B.indptr) and np.array_equal(A.indices, B.indices)
        R = B.tocsr()
        D = (A - B).tocsr()
        num = np.array([np.abs(D.data[D.indptr[i]:D.indptr[i + 1]]).max(initial=0.0) for i in range(D.shape[0])])
        den = np.array([np.abs(R.data[R.indptr[i]:R.indptr[i + 1]]).max() for i in range(R.shape[0])])
        err = float((num / den).max())
        print(f"  {name}: max row error vs high-precision oracle {err:.2e}", flush=True)
        assert err <= 1e-10
    l2 = _poisson_l2(X, Lap.tocsr(), E.tocsr())
    ref_l2 = _poisson_l2(X, sp.csc_matrix((ref["Lap"][2], ref["Lap"][1], ref["Lap"][0]), shape=ref["Lap"][3]).tocsr(),
                         sp.csc_matrix((ref["E"][2], ref["E"][1], ref["E"][0]), shape=ref["E"][3]).tocsr())
    print(f"  relative L2 error {l2:.4e} (oracle operators {ref_l2:.4e})", flush=True)
    assert abs(l2 - ref_l2) <= 1e-8 * ref_l2
    assert l2 < 5e-3
