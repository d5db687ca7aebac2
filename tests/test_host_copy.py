"""Host side of the Python mirror: the thread count of large host copies and
the threaded copy that gives every returned matrix its own index arrays."""
import numpy as np

from rbffd_amd.operators import _host_copy, host_threads


def test_host_threads_follows_the_environment(monkeypatch):
    monkeypatch.setenv("RBFFD_HOST_THREADS", "3")
    monkeypatch.setenv("OMP_NUM_THREADS", "7")
    assert host_threads() == 3
    monkeypatch.setenv("RBFFD_HOST_THREADS", "0")
    assert host_threads() == 7
    monkeypatch.setenv("OMP_NUM_THREADS", "12abc")  # a leading integer, as atoi reads it
    assert host_threads() == 12
    monkeypatch.setenv("OMP_NUM_THREADS", "100")
    assert host_threads() == 32
    monkeypatch.delenv("RBFFD_HOST_THREADS")
    monkeypatch.setenv("OMP_NUM_THREADS", "none")
    assert 1 <= host_threads() <= 32


def test_host_copy_is_an_equal_independent_array(monkeypatch):
    for threads in ("1", "5"):
        monkeypatch.setenv("RBFFD_HOST_THREADS", threads)
        # below and above the size from which the copy is split over threads,
        # with a length that does not divide evenly
        for n in (1000, 3_000_001):
            a = np.random.default_rng(n).integers(0, 1 << 40, size=n, dtype=np.int64)
            b = _host_copy(a)
            assert b.dtype == a.dtype and b.shape == a.shape
            assert np.array_equal(a, b)
            b[0] += 1
            b[-1] += 1
            assert a[0] != b[0] and a[-1] != b[-1]
            # the int64 -> int32 index conversion made before the scipy matrices
            c = _host_copy(a % (1 << 31), np.int32)
            assert c.dtype == np.int32 and np.array_equal(c, a % (1 << 31))
