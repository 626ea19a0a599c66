"""Seeded random-ellipsoid phantoms (no reference counterpart: the reference
loads an external Shepp-Logan volume).

An ellipsoid is a dict with ``center`` (mm, volume frame: x right, y up the
image rows, z along the rotation axis), ``axes`` (semi-axes, mm), ``rot`` (3x3
rotation whose columns are the ellipsoid's axes) and an additive ``density``.
"""
from __future__ import annotations

import numpy as np


def _random_rotation(rng: np.random.Generator) -> np.ndarray:
    # uniform unit quaternion (Shoemake) -> rotation matrix
    u1, u2, u3 = rng.random(3)
    q = np.array([np.sqrt(1 - u1) * np.sin(2 * np.pi * u2), np.sqrt(1 - u1) * np.cos(2 * np.pi * u2),
                  np.sqrt(u1) * np.sin(2 * np.pi * u3), np.sqrt(u1) * np.cos(2 * np.pi * u3)])
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def random_ellipsoids(n: int, seed: int, radius_mm: float, half_height_mm: float,
                      axes_mm=(5.0, 100.0), density=(-0.5, 1.0)) -> list[dict]:
    """``n`` ellipsoids with semi-axes uniform in ``axes_mm`` (scaled down when
    they would not fit), uniform random orientation and density uniform in
    ``density``; each lies inside the cylinder of ``radius_mm`` and
    ``half_height_mm`` (its bounding sphere does)."""
    rng = np.random.default_rng(seed)
    lim = min(radius_mm, half_height_mm)
    out = []
    for _ in range(n):
        axes = rng.uniform(axes_mm[0], axes_mm[1], 3)
        rmax = axes.max()
        if rmax > 0.95 * lim:
            axes *= 0.95 * lim / rmax
            rmax = axes.max()
        free_r = max(radius_mm - rmax, 0.0)
        free_z = max(half_height_mm - rmax, 0.0)
        rr = free_r * np.sqrt(rng.random())
        ph = rng.uniform(0, 2 * np.pi)
        cz = rng.uniform(-free_z, free_z)
        out.append({"center": (rr * np.cos(ph), rr * np.sin(ph), cz),
                    "axes": tuple(float(a) for a in axes),
                    "rot": _random_rotation(rng),
                    "density": float(rng.uniform(density[0], density[1]))})
    return out
