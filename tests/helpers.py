"""Shared helpers for the parity tests: seeded ray batches and image metrics."""
import numpy as np

from noor_amd.renderer import rays_from_numpy


def random_rays(n, center, radius, seed, tmax=None):
    """Rays from points on a sphere around the scene aimed at random interior points."""
    rng = np.random.default_rng(seed)
    u = rng.normal(size=(n, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    o = center + radius * u
    t = center + 0.5 * radius * rng.uniform(-1, 1, (n, 3))
    d = t - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return rays_from_numpy(o.astype(np.float32), d.astype(np.float32), tmax)


def vertex_rays(vertices, n, seed, jitter=0.0):
    """Rays aimed exactly at mesh vertices and edge midpoints (watertightness cases)."""
    rng = np.random.default_rng(seed)
    idx = rng.integers(0, len(vertices), n)
    tgt = vertices[idx, :3].astype(np.float64)
    o = tgt + rng.normal(size=(n, 3)) * 3.0
    d = tgt - o + jitter * rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return rays_from_numpy(o.astype(np.float32), d.astype(np.float32))


def mean_rel_err(a, b, eps=1e-3):
    a = np.asarray(a, np.float64)[..., :3]
    b = np.asarray(b, np.float64)[..., :3]
    return float(np.mean(np.abs(a - b) / (np.abs(b) + eps)))
