"""FP64 CPU oracle of the conduction-velocity fit -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import this module, and only as the checker.

Restates utility/conduction_velocity.py:23-70 (per point: the points within
8 mm by cKDTree.query_ball_point, i included; fewer than 4 -> NaN; the plane
lat = a x + b y + c z + d by np.linalg.lstsq(rcond=None) on the raw
coordinates; g = (a, b, c); v = g / |g|^2, |v| = 1 / |g| when |g|^2 > 1e-8,
else NaN; mean = np.nanmean of the magnitudes) with the same library calls.

Pinned by tests/test_oracle_cv.py: planar activation (the fit is exact, so
v = n * speed for wave direction n), the closed-form minimum-norm solution on
a flat sheet (coplanar points, rank-deficient fit), an independent
normal-equation solve on centred coordinates for curved activation, the
NaN rules (sparse points, constant LAT) and the inclusive radius on lattice
ties.
"""
from __future__ import annotations

import numpy as np
from scipy.spatial import cKDTree

NEIGHBOR_RADIUS_MM = 8.0  # utility/conduction_velocity.py:37


def conduction_velocity(xyz, lat, radius=NEIGHBOR_RADIUS_MM):
    """(vectors (n x 3), magnitudes (n,), mean) of utility/conduction_velocity.py:33-70."""
    xyz = np.asarray(xyz, dtype=np.float64)
    lat = np.asarray(lat)
    tree = cKDTree(xyz)
    n = xyz.shape[0]
    vec = np.zeros((n, 3))
    mag = np.zeros(n)
    for i in range(n):
        idx = tree.query_ball_point(xyz[i], radius)
        if len(idx) < 4:
            vec[i, :] = np.nan
            mag[i] = np.nan
            continue
        pts = xyz[idx]
        A = np.column_stack((pts, np.ones(len(pts))))
        coeffs = np.linalg.lstsq(A, lat[idx], rcond=None)[0]
        g = coeffs[:3]
        gn2 = np.dot(g, g)
        if gn2 > 1e-8:
            vec[i, :] = g / gn2
            mag[i] = 1.0 / np.linalg.norm(g)
        else:
            vec[i, :] = np.nan
            mag[i] = np.nan
    with np.errstate(all="ignore"):
        mean = np.nanmean(mag) if np.isfinite(mag).any() else np.nan
    return vec, mag, mean
