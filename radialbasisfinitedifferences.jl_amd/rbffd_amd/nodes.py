"""Seeded synthetic node sets (jittered Cartesian grids).

Node order is row-major with x fastest.  Edge nodes stay on the grid; every
interior node is moved by an independent uniform offset in [-a h, a h] per
coordinate, drawn with numpy's PCG64 (default_rng(seed)) in node order.
"""
from __future__ import annotations

import numpy as np


def jittered_grid(nx: int, ny: int, jitter: float, seed: int, lx: float = 1.0, ly: float = 1.0,
                  x0: float = 0.0, y0: float = 0.0):
    """(nx*ny, 2) float64 nodes on [x0, x0+lx] x [y0, y0+ly]."""
    hx = lx / (nx - 1)
    hy = ly / (ny - 1)
    ix = np.arange(nx, dtype=np.float64)
    iy = np.arange(ny, dtype=np.float64)
    X = np.empty((ny, nx, 2), np.float64)
    X[:, :, 0] = x0 + ix[None, :] * hx
    X[:, :, 1] = y0 + iy[:, None] * hy
    rng = np.random.default_rng(seed)
    inner = X[1:-1, 1:-1, :]
    off = rng.uniform(-jitter, jitter, size=inner.shape)
    off[..., 0] *= hx
    off[..., 1] *= hy
    X[1:-1, 1:-1, :] = inner + off
    return X.reshape(-1, 2)


def config_nodes(name: str):
    """Node sets of the benchmark configurations, with their RBF-FD parameters."""
    if name == "poisson2500":   # 51 x 51, h = 0.02, r^5, degree 2, n = 13
        return jittered_grid(51, 51, 0.25, seed=1), dict(p=5, polydeg=2, n=13)
    if name == "gen1m":         # 1024 x 1024, r^7, degree 4, n = 31
        return jittered_grid(1024, 1024, 0.25, seed=2), dict(p=7, polydeg=4, n=31)
    if name == "hyper4m":       # 2048 x 2048, +-0.3h, r^9, degree 6, n = 57
        return jittered_grid(2048, 2048, 0.3, seed=3), dict(p=9, polydeg=6, n=57)
    if name == "advdiff4m":     # 2048 x 2048, r^7, degree 4, n = 31
        return jittered_grid(2048, 2048, 0.25, seed=5), dict(p=7, polydeg=4, n=31)
    raise KeyError(name)
