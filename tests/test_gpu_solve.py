"""`D \\ f` of the reference's PDE harness on the GPU (rbffd_lsq_solve_device):
the oversampled Poisson system of test/poisson_test.jl:74-121 built from GPU
operators and solved by the direct (QR) and iterative (CGLS) methods, the
square Dirichlet Poisson system of the 51 x 51 configuration, and small systems
against numpy's least-squares routine (LAPACK gelsd), the textbook answer
both methods must reproduce."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as sla

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def poisson(ctx, O, fitted_nodes):
    import rbffd_amd
    from test_oracle_weights import load_poisson_case, poisson_system

    X, Y, yin, yneu, ydir, xn, yn = load_poisson_case(O, fitted_nodes)
    E, Dx, Dy, Dxx, Dyy, _ = rbffd_amd.generate_operator(X, Y, 3, 20, 3, ctx=ctx)
    A, b = poisson_system([E, Dx, Dy, Dxx, Dyy], X, Y, yin, yneu, ydir, xn, yn, O)
    direct = np.linalg.lstsq(A.toarray(), b, rcond=None)[0]
    return A, b, E, Y, direct


def test_poisson_fixture_qr(ctx, poisson):
    import rbffd_amd
    from test_oracle_weights import poisson_rel_error

    A, b, E, Y, direct = poisson
    print(f"Poisson fixture, QR: D {A.shape}, nnz {A.nnz}", flush=True)
    x = rbffd_amd.solve(A, b, method="qr", ctx=ctx)
    d = _rel(x, direct)
    err, err_ref = poisson_rel_error(E, x, Y), poisson_rel_error(E, direct, Y)
    print(f"  |x - lstsq| / |lstsq| = {d:.2e}; error {err:.10f} (lstsq {err_ref:.10f})", flush=True)
    assert d <= 1e-10
    assert abs(err - err_ref) <= 1e-10
    assert err < 0.0027   # test/poisson_test.jl:132


def test_poisson_fixture_cgls(ctx, poisson):
    import rbffd_amd
    from test_oracle_weights import poisson_rel_error

    A, b, E, Y, direct = poisson
    print("Poisson fixture, CGLS to 1e-12", flush=True)
    x, (its, rel) = rbffd_amd.solve(A, b, method="cgls", tol=1e-12, max_iters=20000, ctx=ctx, info=True)
    d = _rel(x, direct)
    err, err_ref = poisson_rel_error(E, x, Y), poisson_rel_error(E, direct, Y)
    print(f"  {its} iterations, relres {rel:.2e}; |x - lstsq| / |lstsq| = {d:.2e}; error {err:.10f}", flush=True)
    assert rel <= 1e-12 and its < 20000
    assert d <= 1e-9
    assert abs(err - err_ref) <= 1e-9
    assert err < 0.0027   # test/poisson_test.jl:132
    # the sums run in a fixed order: a second solve gives the same bits
    x2, (its2, rel2) = rbffd_amd.solve(A, b, method="cgls", tol=1e-12, max_iters=20000, ctx=ctx, info=True)
    assert its2 == its and rel2 == rel and np.array_equal(x, x2)


def test_square_poisson_51x51_qr(ctx):
    """Square Dirichlet Poisson system (51 x 51 nodes, r^5, degree 2, n = 13):
    QR against the sparse LU of scipy (what `\\` runs for a square D)."""
    import rbffd_amd
    from rbffd_amd.nodes import config_nodes

    X, prm = config_nodes("poisson2500")
    E, Lap = rbffd_amd.operators(X, X, prm["p"], prm["n"], prm["polydeg"], ["E", "Lap"], ctx=ctx)
    x, y = X[:, 0], X[:, 1]
    u = np.sin(np.pi * x) * np.sin(np.pi * y)
    edge = np.where((x <= 0) | (x >= 1) | (y <= 0) | (y >= 1))[0]
    A = sp.lil_matrix(Lap.tocsr())
    A[edge, :] = E.tocsr()[edge, :]
    A = A.tocsr()
    rhs = -2 * np.pi**2 * u
    rhs[edge] = u[edge]
    ref = sla.spsolve(A.tocsc(), rhs)
    sol = rbffd_amd.solve(A, rhs, method="auto", ctx=ctx)
    d = _rel(sol, ref)
    print(f"51 x 51 square system: |x - splu| / |splu| = {d:.2e}, L2 error {_rel(sol, u):.4e}", flush=True)
    assert d <= 1e-10
    assert _rel(sol, u) < 5e-3


@pytest.mark.parametrize("M,N,method", [(60, 20, "qr"), (60, 20, "cgls"), (40, 40, "qr"), (40, 40, "cgls"),
                                        (300, 7, "qr"), (300, 7, "cgls")])
def test_small_systems_match_lstsq(ctx, M, N, method):
    import rbffd_amd

    rng = np.random.default_rng(M * 1000 + N)
    A = sp.random(M, N, density=0.3, random_state=rng, format="csr") + sp.eye(M, N, format="csr") * 3.0
    b = rng.standard_normal(M)
    ref = np.linalg.lstsq(A.toarray(), b, rcond=None)[0]
    x, (its, rel) = rbffd_amd.solve(A, b, method=method, tol=1e-13, max_iters=5000, ctx=ctx, info=True)
    d = _rel(x, ref)
    print(f"{M} x {N} {method}: {its} iterations, |x - lstsq| / |lstsq| = {d:.2e}", flush=True)
    assert d <= 1e-10


def test_cgls_iteration_budget_and_start(ctx):
    """max_iters is honoured exactly (37 = one recorded block of 32 plus 5
    single iterations), a start value is used (r = b - D x0), and a CSR whose
    rowptr does not start at 0 is refused."""
    import rbffd_amd
    from rbffd_amd.operators import Engine

    rng = np.random.default_rng(7)
    M, N = 400, 150
    A = (sp.random(M, N, density=0.05, random_state=rng, format="csr") + sp.eye(M, N, format="csr")).tocsr()
    A.sort_indices()
    b = rng.standard_normal(M)
    _, (its, rel) = rbffd_amd.solve(A, b, method="cgls", tol=0.0, max_iters=37, ctx=ctx, info=True)
    assert its == 37 and rel > 0.0
    ref = np.linalg.lstsq(A.toarray(), b, rcond=None)[0]
    eng = Engine(ctx)
    x0 = ref + 0.1 * rng.standard_normal(N)
    dev = [ctx.to_device(a) for a in (A.indptr.astype(np.int64), A.indices.astype(np.int32), A.data, b, x0)]
    its, rel = eng.lsq_solve(*dev, method="cgls", tol=1e-13, max_iters=2000)
    assert _rel(dev[4].get(), ref) <= 1e-10
    bad = ctx.to_device(A.indptr.astype(np.int64) + 1)
    with pytest.raises(ValueError):
        eng.lsq_solve(bad, *dev[1:], method="cgls")


def test_rank_deficient_and_bad_arguments(ctx):
    import rbffd_amd

    rng = np.random.default_rng(3)
    A = (sp.random(30, 10, density=0.4, random_state=rng, format="csr") + sp.eye(30, 10, format="csr")).tolil()
    A[:, 4] = 0.0   # an empty column
    A = A.tocsr()
    A.eliminate_zeros()
    b = rng.standard_normal(30)
    with pytest.raises(rbffd_amd.SingularException):
        rbffd_amd.solve(A, b, method="qr", ctx=ctx)
    # CGLS leaves the free unknown at its start (0) and solves for the rest
    x = rbffd_amd.solve(A, b, method="cgls", tol=1e-13, max_iters=2000, ctx=ctx)
    keep = np.r_[0:4, 5:10]
    ref = np.linalg.lstsq(A.toarray()[:, keep], b, rcond=None)[0]
    assert x[4] == 0.0 and _rel(x[keep], ref) <= 1e-10
    with pytest.raises(ValueError):
        rbffd_amd.solve(sp.random(5, 9, density=0.5, format="csr"), np.ones(5), ctx=ctx)   # M < N
    with pytest.raises(rbffd_amd.NonFiniteError):
        rbffd_amd.solve(sp.eye(6, format="csr") * np.nan, np.ones(6), method="cgls", ctx=ctx)
