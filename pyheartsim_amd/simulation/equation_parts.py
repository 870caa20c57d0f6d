"""Per-voxel coefficient packing (drop-in for simulation/equation_parts.py).

``load_fiber`` and ``heart_model_equation_parts`` keep the reference's
signatures and the N x 21 ``P_2d`` layout of simulation/equation_parts.py:36-114:

  cols 0-5   4 * delta_k * D_ii   (delta_k = 1 when face neighbour k exists)
  cols 6-14  anisotropy cross terms (identically 0 for D = I)
  cols 15-19 model parameters  (MS: tau_open, tau_close, tau_in, tau_out, v_gate;
                                AP: k, a, epsilon_0, mu1, mu2)
  col 20     conductivity c

``load_fiber`` returns an ``IsotropicFibers`` sequence: it indexes like the
reference's list of N identity matrices (fibres are disabled there,
equation_parts.py:19) without materialising N 3x3 arrays, and lets the
device path recognise D = I without scanning N matrices.
"""
from __future__ import annotations

import numpy as np


class IsotropicFibers:
    """Sequence of N identity diffusion tensors (reference D0 with fibres off)."""

    def __init__(self, n_voxel: int):
        self.n = int(n_voxel)

    def __len__(self):
        return self.n

    def __getitem__(self, i):
        if isinstance(i, slice):
            return [np.eye(3) for _ in range(*i.indices(self.n))]
        if i < -self.n or i >= self.n:
            raise IndexError(i)
        return np.eye(3)

    def __iter__(self):
        for _ in range(self.n):
            yield np.eye(3)


def load_fiber(n_voxel):
    """Reference equation_parts.py:17-34 (fiber_flag = 0 -> D = I everywhere)."""
    return IsotropicFibers(n_voxel)


def tensor_components(D0, n_voxel):
    """(9, N) float64 array D_ij for the N diffusion tensors."""
    if isinstance(D0, IsotropicFibers):
        out = np.zeros((9, n_voxel))
        out[0] = out[4] = out[8] = 1.0
        return out
    arr = np.asarray([np.asarray(d, dtype=np.float64).reshape(9) for d in D0])
    return np.ascontiguousarray(arr.T)


def heart_model_equation_parts(simulation_parameters, n_voxel, D0, neighbor_id_2d, parameter):
    """Reference equation_parts.py:36-114 (same N x 21 float64 output)."""
    D = tensor_components(D0, n_voxel)
    D11, D12, D13, D21, D22, D23, D31, D32, D33 = D
    nb = np.asarray(neighbor_id_2d)
    delta = (nb >= 0).astype(np.float64)
    nb0 = np.where(nb < 0, 0, nb)
    P = np.zeros((n_voxel, 21))
    P[:, 0] = 4 * delta[:, 0] * D11
    P[:, 1] = 4 * delta[:, 1] * D11
    P[:, 2] = 4 * delta[:, 2] * D22
    P[:, 3] = 4 * delta[:, 3] * D22
    P[:, 4] = 4 * delta[:, 4] * D33
    P[:, 5] = 4 * delta[:, 5] * D33

    def pair(k):
        return delta[:, k] * delta[:, k + 1]

    def diff(Dc, k):
        return Dc[nb0[:, k]] - Dc[nb0[:, k + 1]]

    P[:, 6] = pair(0) * (pair(0) * diff(D11, 0) + pair(2) * diff(D21, 2) + pair(4) * diff(D31, 4))
    P[:, 7] = pair(2) * (pair(0) * diff(D12, 0) + pair(2) * diff(D22, 2) + pair(4) * diff(D32, 4))
    P[:, 8] = pair(4) * (pair(0) * diff(D13, 0) + pair(2) * diff(D23, 2) + pair(4) * diff(D33, 4))
    P[:, 9] = 2 * delta[:, 6] * delta[:, 8] * D12
    P[:, 10] = 2 * delta[:, 7] * delta[:, 9] * D12
    P[:, 11] = 2 * delta[:, 14] * delta[:, 16] * D13
    P[:, 12] = 2 * delta[:, 15] * delta[:, 17] * D13
    P[:, 13] = 2 * delta[:, 10] * delta[:, 12] * D23
    P[:, 14] = 2 * delta[:, 11] * delta[:, 13] * D23
    if simulation_parameters["heart_model_flag"] == 0:
        keys = ("tau_open_voxel", "tau_close_voxel", "tau_in_voxel", "tau_out_voxel", "v_gate_voxel")
    else:
        keys = ("k_voxel", "a_voxel", "epsilon_0_voxel", "mu1_voxel", "mu2_voxel")
    for col, key in zip(range(15, 20), keys):
        P[:, col] = parameter[key]
    P[:, 20] = parameter["c_voxel"]
    return P
