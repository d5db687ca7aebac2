"""The oracle's CSR product and CSR assembly against scipy.sparse (an
independent implementation): y = A x for one and several node-major fields,
ragged rows, empty rows and explicit zeros."""
import numpy as np
import pytest
import scipy.sparse as sp


def _random_csr(rng, M, N, max_row):
    lens = rng.integers(0, max_row + 1, size=M)
    lens[::7] = 0                                    # empty rows
    rowptr = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    col = rng.integers(0, N, size=rowptr[-1]).astype(np.int32)
    val = rng.standard_normal(rowptr[-1])
    val[::5] = 0.0                                   # explicit zeros
    return rowptr, col, val


@pytest.mark.parametrize("F", [1, 2, 8])
def test_spmv_csr_matches_scipy(O, F):
    rng = np.random.default_rng(F)
    M, N = 300, 250
    rowptr, col, val = _random_csr(rng, M, N, 12)
    x = rng.standard_normal((N, F)) if F > 1 else rng.standard_normal(N)
    y = O.spmv_csr(rowptr, col, val, x)
    A = sp.csr_matrix((val, col, rowptr), shape=(M, N))
    ref = A @ x
    assert y.shape == ref.shape
    assert np.allclose(y, ref, rtol=1e-13, atol=1e-13)


def test_csr_from_rows_is_the_sorted_stencil_layout(O):
    rng = np.random.default_rng(3)
    M, n, N = 40, 7, 90
    st = np.stack([rng.choice(N, n, replace=False) for _ in range(M)]).astype(np.int32)
    W = rng.standard_normal((M, n))
    rowptr, col, val = O.csr_from_rows(st, W)
    assert np.array_equal(rowptr, np.arange(M + 1) * n)
    col = np.asarray(col).reshape(M, n)
    val = np.asarray(val).reshape(M, n)
    assert np.all(np.diff(col, axis=1) > 0)
    dense = np.zeros((M, N))
    np.put_along_axis(dense, st.astype(np.int64), W, 1)
    rebuilt = np.zeros((M, N))
    np.put_along_axis(rebuilt, col.astype(np.int64), val, 1)
    assert np.array_equal(dense, rebuilt)
