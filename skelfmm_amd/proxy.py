"""Proxy surfaces used by the interpolative decompositions of the build.

The reference takes these from the absent ``simpletree`` package
(``skelfmm/util.py:6-7,72-78``): ``get_cube_surface(center, side, nproxy=)`` and
``get_sphere_surface(center, diameter, npoints=)``. Their point layouts are not
recoverable, so this module defines its own:

* cube, 2D: ``ceil(nproxy/4)`` points per side at the side's cell midpoints;
* cube, 3D: an ``m x m`` grid of face-cell midpoints on each of the six faces,
  ``m = ceil(sqrt(nproxy/6))``;
* sphere (3D only): a Fibonacci lattice of ``npoints`` points on the sphere of
  the given diameter.

Corner and edge points are never duplicated, so the proxy rows are distinct.
"""

from __future__ import annotations

import math

import numpy as np

__all__ = ["get_cube_surface", "get_sphere_surface", "sphere_proxy_normals"]


def get_cube_surface(center, side, nproxy):
    center = np.asarray(center, dtype=np.float64)
    d = center.shape[0]
    h = 0.5 * side
    if d == 2:
        m = max(1, int(math.ceil(nproxy / 4)))
        t = (np.arange(m) + 0.5) / m * side - h
        pts = [
            np.stack([t, np.full(m, -h)], axis=1),
            np.stack([np.full(m, h), t], axis=1),
            np.stack([-t, np.full(m, h)], axis=1),
            np.stack([np.full(m, -h), -t], axis=1),
        ]
        return np.vstack(pts) + center[None, :]
    if d == 3:
        m = max(1, int(math.ceil(math.sqrt(nproxy / 6.0))))
        t = (np.arange(m) + 0.5) / m * side - h
        a, b = np.meshgrid(t, t, indexing="ij")
        a = a.ravel()
        b = b.ravel()
        faces = []
        for axis in range(3):
            for sgn in (-1.0, 1.0):
                f = np.empty((a.shape[0], 3))
                others = [k for k in range(3) if k != axis]
                f[:, axis] = sgn * h
                f[:, others[0]] = a
                f[:, others[1]] = b
                faces.append(f)
        return np.vstack(faces) + center[None, :]
    raise ValueError("cube surface needs a 2D or 3D center")


def get_sphere_surface(center, diameter, npoints):
    center = np.asarray(center, dtype=np.float64)
    if center.shape[0] != 3:
        raise ValueError("sphere surface is 3D only")
    n = max(1, int(npoints))
    k = np.arange(n, dtype=np.float64)
    z = 1.0 - (2.0 * k + 1.0) / n
    rho = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    phi = k * (math.pi * (3.0 - math.sqrt(5.0)))
    pts = np.stack([rho * np.cos(phi), rho * np.sin(phi), z], axis=1)
    return 0.5 * diameter * pts + center[None, :]


def sphere_proxy_normals(proxy_points, box_center):
    """Outward unit normals of proxy points on a sphere (``util.py:66-69``)."""
    n = proxy_points - np.asarray(box_center)[None, :]
    return n / np.linalg.norm(n, axis=1, keepdims=True)
