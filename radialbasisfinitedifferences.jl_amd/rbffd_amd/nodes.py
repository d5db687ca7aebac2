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


def ghosted_rectangle(nodes: int, aspect: float, jitter: float, seed: int):
    """A jittered grid on [0, aspect] x [0, 1] with about `nodes` points, plus
    one ghost node per boundary node at distance h along the outward normal
    (corners take the normal of the edge they are listed under).  Returns
    (X, X_idx_in, X_idx_bc, X_idx_bc_g) with 0-based indices and the four
    markers left, right, bottom, top."""
    ny = max(4, int(round(np.sqrt(nodes / aspect))))
    nx = max(4, int(round(nodes / ny)))
    X = jittered_grid(nx, ny, jitter, seed, lx=aspect, ly=1.0)
    hx, hy = aspect / (nx - 1), 1.0 / (ny - 1)
    ix = np.tile(np.arange(nx), ny)
    iy = np.repeat(np.arange(ny), nx)
    idx = np.arange(nx * ny)
    left = idx[ix == 0]
    right = idx[ix == nx - 1]
    bottom = idx[(iy == 0) & (ix > 0) & (ix < nx - 1)]
    top = idx[(iy == ny - 1) & (ix > 0) & (ix < nx - 1)]
    inner = idx[(ix > 0) & (ix < nx - 1) & (iy > 0) & (iy < ny - 1)]
    bcs, ghosts, extra = [], [], []
    nxt = nx * ny
    for b, normal in ((left, (-hx, 0.0)), (right, (hx, 0.0)), (bottom, (0.0, -hy)), (top, (0.0, hy))):
        g = X[b] + np.array(normal)
        bcs.append(b)
        ghosts.append(np.arange(nxt, nxt + len(b)))
        extra.append(g)
        nxt += len(b)
    X = np.concatenate([X] + extra)
    return X, inner, bcs, ghosts


def batch_sets(count: int = 64, seed: int = 4, lo: float = 1e4, hi: float = 5e5):
    """The uneven batch: `count` ghosted rectangles, sizes log-uniform in
    [lo, hi], aspect ratio U(1, 4), polynomial degree from {2, 3, 4, 5} with
    stencil n = 2 C(deg + 2, 2) + 1 and PHS power m = 2 floor(deg / 2) + 3.
    Returns a list of dicts (X, X_idx_in, X_idx_bc, X_idx_bc_g, p, n, polydeg)."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        nodes = float(np.exp(rng.uniform(np.log(lo), np.log(hi))))
        aspect = float(rng.uniform(1.0, 4.0))
        deg = int(rng.choice([2, 3, 4, 5]))
        n = 2 * (deg + 1) * (deg + 2) // 2 + 1
        X, inner, bcs, ghosts = ghosted_rectangle(int(nodes), aspect, 0.25, seed * 1000 + i)
        out.append(dict(X=X, X_idx_in=inner, X_idx_bc=bcs, X_idx_bc_g=ghosts, p=2 * (deg // 2) + 3, n=n,
                        polydeg=deg))
    return out


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
