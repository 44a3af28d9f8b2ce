"""Deterministic ray sets for traversal parity tests (rays[n, 8] = origin, direction, tMin, tMax)."""
from __future__ import annotations

import numpy as np


def random_rays(scene, n: int, seed: int = 0, spread: float = 0.1) -> np.ndarray:
    """Origins inside the scene bounds (slightly enlarged), isotropic directions."""
    rng = np.random.default_rng(seed)
    lo = scene.positions.min(0)
    hi = scene.positions.max(0)
    ext = hi - lo
    o = rng.uniform(lo - spread * ext, hi + spread * ext, (n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.zeros((n, 8), np.float32)
    r[:, :3] = o
    r[:, 3:6] = d
    r[:, 6] = 1e-4
    r[:, 7] = np.inf
    # a few axis-aligned and near-axis directions (zero / tiny components exercise safe(1e-9))
    k = min(n, 64)
    axes = np.eye(3)[rng.integers(0, 3, k)] * rng.choice([-1.0, 1.0], (k, 1))
    axes[k // 2:] += rng.normal(scale=1e-11, size=(k - k // 2, 3))
    r[:k, 3:6] = axes / np.linalg.norm(axes, axis=1, keepdims=True)
    return r


def camera_rays(cam, n: int, seed: int = 0) -> np.ndarray:
    """Pinhole camera rays (Camera::generateRay without lens), computed in float32 numpy."""
    rng = np.random.default_rng(seed)
    W, H = cam.width, cam.height
    px = rng.integers(0, W, n)
    py = rng.integers(0, H, n)
    jx, jy = rng.uniform(0, 1, n), rng.uniform(0, 1, n)
    s = ((px + jx) / W).astype(np.float32)
    t = ((H - 1 - py + jy) / H).astype(np.float32)
    llc = np.array(cam.lower_left, np.float32)
    hor = np.array(cam.horizontal, np.float32)
    ver = np.array(cam.vertical, np.float32)
    org = np.array(cam.origin, np.float32)
    tgt = llc + hor * s[:, None] + ver * t[:, None]
    d = tgt - org
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.zeros((n, 8), np.float32)
    r[:, :3] = org
    r[:, 3:6] = d
    r[:, 6] = 1e-4
    r[:, 7] = np.inf
    return r
