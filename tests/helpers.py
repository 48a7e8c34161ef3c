"""Shared helpers of the test-suite (random states, comparisons)."""
from __future__ import annotations

import numpy as np


def random_motion(st: dict, sc, seed: int, vscale=50.0, wscale=200.0, quats=True):
    """Give the moving particles random velocities, angular velocities and
    (with quats) random orientation quaternions. Vertices are left as they are."""
    rng = np.random.Generator(np.random.PCG64(seed))
    W, N = sc.n_walls, sc.n_rows
    n = N - W
    st["v"][W:] = rng.normal(size=(n, 3)) * vscale
    st["v12"][W:] = st["v"][W:] + rng.normal(size=(n, 3))
    st["w"][W:] = rng.normal(size=(n, 3)) * wscale
    st["wbf"][W:] = rng.normal(size=(n, 3)) * wscale
    st["w12bf"][W:] = st["wbf"][W:] + rng.normal(size=(n, 3))
    if quats:
        for k in ("q", "q12", "qall"):
            q = rng.normal(size=(n, 4))
            q /= np.linalg.norm(q, axis=1, keepdims=True)
            st[k][W:] = q
    return st


def rel_err(a, b, floor=1e-300):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.maximum(np.abs(b), floor)
    return float(np.max(np.abs(a - b) / den)) if a.size else 0.0


def max_rel(a, b, scale=None):
    """max |a-b| / max(|b|, scale) with scale defaulting to max|b| * 1e-3."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    if a.size == 0:
        return 0.0
    s = scale if scale is not None else max(float(np.max(np.abs(b))) * 1e-3, 1e-300)
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), s)))
