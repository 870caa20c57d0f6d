"""Host helpers of the unipolar electrogram (drop-in for
simulation/unipolar_electrogram_equation_parts.py, same signatures).

The device path does not use the N x E distance arrays or the T x N gradient
arrays these helpers build: it folds the gradient stencil into a per-electrode
lead field on the device.  They are kept for callers of the reference API.
"""
from __future__ import annotations

import numpy as np

from .equation_parts import tensor_components


def extract_diffusion_tensor_components(D0, n_voxel, dtype=np.float64):
    """Reference :17-39 -> nine length-N arrays D11 .. D33."""
    D = tensor_components(D0, n_voxel).astype(dtype)
    return tuple(D[i].copy() for i in range(9))


def compute_electrode_distances(electrode_xyz, voxel, dtype=np.float64):
    """Reference :41-58 -> l, l_x, l_y, l_z (N x E); l clamped to >= 1 mm."""
    electrode_xyz = np.asarray(electrode_xyz, dtype=np.float64)
    voxel = np.asarray(voxel, dtype=np.float64)
    l_x = (voxel[:, 0:1] - electrode_xyz[None, :, 0]).astype(dtype)
    l_y = (voxel[:, 1:2] - electrode_xyz[None, :, 1]).astype(dtype)
    l_z = (voxel[:, 2:3] - electrode_xyz[None, :, 2]).astype(dtype)
    l = np.sqrt(l_x ** 2 + l_y ** 2 + l_z ** 2)
    l[l < 1] = 1
    return l, l_x, l_y, l_z


def compute_gradients(action_potential, neighbor_id_2d, Delta):
    """Reference :60-103 -> central differences, 0 where an axis neighbour is missing."""
    nb = np.asarray(neighbor_id_2d)[:, :6]
    ap = np.asarray(action_potential)
    out = []
    for axis in range(3):
        p, m = nb[:, 2 * axis], nb[:, 2 * axis + 1]
        ok = (p >= 0) & (m >= 0)
        g = (ap[:, np.where(p < 0, 0, p)] - ap[:, np.where(m < 0, 0, m)]) / (2 * Delta)
        g[:, ~ok] = 0
        out.append(g)
    return tuple(out)


def broadcast_diffusion_tensors(D11, D12, D13, D21, D22, D23, D31, D32, D33, n_electrode,
                                dtype=np.float64):
    """Reference :105-116."""
    return tuple(np.tile(np.asarray(d)[:, None], (1, n_electrode)).astype(dtype)
                 for d in (D11, D12, D13, D21, D22, D23, D31, D32, D33))


def normalize_electrogram_magnitude(electrogram_unipolar, typical_magnitude=1.0):
    """Reference :118-124: scale so the mean peak-to-peak over electrodes is 1."""
    egm = np.asarray(electrogram_unipolar)
    voltage = np.ptp(egm, axis=0)
    return egm * (typical_magnitude / np.mean(voltage))
