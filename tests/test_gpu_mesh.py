"""The reference's mesh Poisson test (test/mesh_import_test.jl:17-158) with
GPU operators and the GPU solve of `D \\ f`: the CGNS mesh fixture
test/data/tominec_Y.cgns imported by rbffd_amd.mesh, r^3, degree 5, n = 42,
err < 0.001, and the same error as with the oracle's operators and LAPACK."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_mesh_poisson_on_gpu(ctx, O, fitted_nodes):
    import rbffd_amd
    from test_mesh import _mesh_case
    from test_oracle_weights import _csc, poisson_error, poisson_rel_error, poisson_system

    X, Y, yin, yneu, ydir, xn, yn = _mesh_case(O, fitted_nodes)
    print(f"mesh Poisson: {len(X)} X nodes, {len(Y)} Y nodes, r^3, degree 5, n = 42", flush=True)
    E, Dx, Dy, Dxx, Dyy, _ = rbffd_amd.generate_operator(X, Y, 3, 42, 5, ctx=ctx)
    A, b = poisson_system([E, Dx, Dy, Dxx, Dyy], X, Y, yin, yneu, ydir, xn, yn, O)
    sol = rbffd_amd.solve(A, b, method="qr", ctx=ctx)
    err = poisson_rel_error(E, sol, Y)
    ref = O.generate_operator(X, Y, 3, 42, 5, ops=("E", "Dx", "Dy", "Dxx", "Dyy"))
    err_ref = poisson_error([_csc(ref[k]) for k in ("E", "Dx", "Dy", "Dxx", "Dyy")], X, Y, yin, yneu, ydir, xn, yn, O)
    print(f"  error {err:.8f} (oracle operators + LAPACK: {err_ref:.8f})", flush=True)
    assert err < 0.001   # test/mesh_import_test.jl:158
    assert abs(err - err_ref) <= 1e-6 * err_ref
