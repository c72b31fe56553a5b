"""Synthetic inputs of the benchmark configurations (seeded, float32).

Every generator is deterministic in its seed (numpy PCG64).  Geometry is
computed in float64 and rounded once to float32.
"""
from __future__ import annotations

import numpy as np


def triangle_soup(n: int, seed: int = 3, mu: float = -6.0, sigma: float = 0.5) -> np.ndarray:
    """LBVH workload: n triangles, centroids U[0,1]^3, size lognormal(mu, sigma),
    uniformly random orientation.  Returns (n, 9) float32."""
    rng = np.random.Generator(np.random.PCG64(seed))
    c = rng.random((n, 3))
    s = rng.lognormal(mu, sigma, n)
    # uniform random rotations from unit quaternions
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q.T
    rot = np.empty((n, 3, 3))
    rot[:, 0, 0] = 1 - 2 * (y * y + z * z)
    rot[:, 0, 1] = 2 * (x * y - z * w)
    rot[:, 0, 2] = 2 * (x * z + y * w)
    rot[:, 1, 0] = 2 * (x * y + z * w)
    rot[:, 1, 1] = 1 - 2 * (x * x + z * z)
    rot[:, 1, 2] = 2 * (y * z - x * w)
    rot[:, 2, 0] = 2 * (x * z - y * w)
    rot[:, 2, 1] = 2 * (y * z + x * w)
    rot[:, 2, 2] = 1 - 2 * (x * x + y * y)
    # equilateral triangle with unit edges, centred at the origin, in the xy-plane
    r = 1.0 / np.sqrt(3.0)
    base = np.array([[r, 0.0, 0.0],
                     [-0.5 * r, 0.5, 0.0],
                     [-0.5 * r, -0.5, 0.0]])
    v = np.einsum("nij,kj->nki", rot, base) * s[:, None, None] + c[:, None, :]
    return v.reshape(n, 9).astype(np.float32)
