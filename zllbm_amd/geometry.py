"""Immersed-boundary marker sets.

Marker placement restated from the reference's default shapes; coordinates are in the node
frame (node index i of level l sits at i * 2**-l), areas as set by the IB shapes' UpdateShape.
"""
from __future__ import annotations

import math

import numpy as np

K_EPS = 1e-6   # kEps, defs_libs.h:103


def circle_markers(center, radius: float, dx: float):
    """GeoShapeDefaultCircle2D::InitialShape (criterion/geometry_default_shape_2d.cpp:23-45):
    n = int(2 pi r / dx + eps) + 1 points at angles 2 pi i / n; arc length per marker 2 pi r / n
    (GeoShapeIBCircle2D::UpdateShape, immersed_boundary/geometry_ib_shape_2d.cpp:41-55)."""
    n = int(2 * math.pi * radius / dx + K_EPS) + 1
    i = np.arange(n, dtype=np.float64) / n
    xy = np.stack([radius * np.cos(2 * math.pi * i) + center[0],
                   radius * np.sin(2 * math.pi * i) + center[1]], axis=1)
    area = np.full(n, 2 * math.pi * radius / n)
    return xy, area


def fibonacci_sphere_markers(center, radius: float, dx: float):
    """GeoShapeDefaultSphere3D::InitialShape method 1 (criterion/geometry_default_shape_3d.cpp:297-318):
    n = int(4 pi r^2 / dx^2 + eps) + 1, x_i = 1 - 2 i / (n - 1), azimuth pi (sqrt 5 - 1) i; area per
    marker 4 pi r^2 / n (GeoShapeIBSphere3D::UpdateShape, geometry_ib_shape_3d.cpp:46-61)."""
    n = int(4 * math.pi * radius * radius / (dx * dx) + K_EPS) + 1
    phi = math.pi * (math.sqrt(5.0) - 1.0)
    i = np.arange(n, dtype=np.float64)
    ux = 1.0 - (i / (n - 1)) * 2
    ur = np.sqrt(1.0 - ux * ux)
    th = phi * i
    xyz = np.stack([ux * radius + center[0], np.cos(th) * ur * radius + center[1],
                    np.sin(th) * ur * radius + center[2]], axis=1)
    area = np.full(n, 4 * math.pi * radius * radius / n)
    return xyz, area
