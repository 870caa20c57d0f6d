"""Neighbour table of a voxel set (drop-in for utility/voxelization.py:102-133).

``find_neighbor_voxel_ids(voxel)`` keeps the reference's definition: for each voxel
and each of the 18 unit offsets (column order +x, -x, +y, -y, +z, -z, then the 12
edge neighbours), the neighbour is the voxel lying within 0.1 of voxel + offset, -1
if none.  The reference answers each of the 18 N queries with a KD-tree; here every
voxel is hashed once by its rounded position relative to the set's minimum corner
and each query is one hash lookup plus the 0.1 distance check, O(N) in NumPy.  The
offsets are unit offsets whatever the voxel spacing, as in the reference.
"""
from __future__ import annotations

import numpy as np

from ..geometry import NEIGHBOR_OFFSETS

TOLERANCE = 0.1


def find_neighbor_voxel_ids(voxel) -> np.ndarray:
    v = np.asarray(voxel, dtype=np.float64)
    n = v.shape[0]
    out = np.full((n, 18), -1, dtype=np.int64)
    if n == 0:
        return out
    origin = v.min(axis=0)
    key = np.rint(v - origin).astype(np.int64)
    lo = key.min(axis=0) - 2
    dims = key.max(axis=0) - lo + 3
    flat = ((key[:, 0] - lo[0]) * dims[1] + (key[:, 1] - lo[1])) * dims[2] + (key[:, 2] - lo[2])
    table = np.full(int(dims[0] * dims[1] * dims[2]), -1, dtype=np.int64)
    # voxels sharing a lattice site: the lowest id keeps it (a KD-tree would return the nearest)
    order = np.argsort(flat, kind="stable")[::-1]
    table[flat[order]] = order
    for c in range(18):
        p = v + NEIGHBOR_OFFSETS[c]
        k = np.rint(p - origin).astype(np.int64)
        inside = np.all((k - lo >= 0) & (k - lo < dims), axis=1)
        cand = np.full(n, -1, dtype=np.int64)
        kf = ((k[inside, 0] - lo[0]) * dims[1] + (k[inside, 1] - lo[1])) * dims[2] + (k[inside, 2] - lo[2])
        cand[inside] = table[kf]
        ok = cand >= 0
        d = np.full(n, np.inf)
        d[ok] = np.linalg.norm(v[cand[ok]] - p[ok], axis=1)
        hit = ok & (d < TOLERANCE)
        out[hit, c] = cand[hit]
    return out
