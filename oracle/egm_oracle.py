"""FP64 CPU oracle of the unipolar electrogram -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import this module, and only as the checker.

Restates, in FP64:
  * electrode distances   simulation/unipolar_electrogram_equation_parts.py:41-58
                          (l clamped to >= 1 mm; components not clamped)
  * gradients             simulation/unipolar_electrogram_equation_parts.py:60-103
                          (central differences, 0 when an axis neighbour is missing)
  * per-voxel term        legacy/compute_unipolar_electrogram_cpu.py:22-32
                          = simulation/unipolar_electrogram_gpu.py:54-69:
                          c_v / l^3 * ((D grad u) . l_vec), summed over voxels
  * normalisation         simulation/unipolar_electrogram_equation_parts.py:118-124
  * electrode LAT         utility/lat_map.py:20-27
Layout follows the GPU API (simulation/unipolar_electrogram_gpu.py:71-152):
action_potential is (T x N), the result is (T x E).

Pinned by tests/test_oracle_egm.py: a one-voxel dipole closed form, the
zero gradient on one-sided axes, linearity, the normalisation invariant,
a brute-force triple loop on tiny inputs, and hand-built electrode signals
with known steepest-downstroke indices (ties, shift to 0).
"""
from __future__ import annotations

import numpy as np


def distances(electrode_xyz, voxel):
    lx = voxel[:, 0:1] - electrode_xyz[None, :, 0]
    ly = voxel[:, 1:2] - electrode_xyz[None, :, 1]
    lz = voxel[:, 2:3] - electrode_xyz[None, :, 2]
    l = np.sqrt(lx ** 2 + ly ** 2 + lz ** 2)
    l[l < 1] = 1
    return l, lx, ly, lz


def gradients(ap, neighbor_id_2d, Delta):
    nb = np.asarray(neighbor_id_2d)[:, :6]
    out = []
    for axis in range(3):
        p, m = nb[:, 2 * axis], nb[:, 2 * axis + 1]
        g = (ap[:, np.where(p < 0, 0, p)] - ap[:, np.where(m < 0, 0, m)]) / (2 * Delta)
        g[:, (p < 0) | (m < 0)] = 0
        out.append(g)
    return out


def electrogram_raw(electrode_xyz, voxel, D, c_voxel, action_potential, Delta, neighbor_id_2d,
                    t_chunk=256):
    """(T x E) un-normalised electrogram in FP64.

    ``D`` is (9, N) row-major tensor components (D11, D12, ..., D33).
    """
    electrode_xyz = np.asarray(electrode_xyz, dtype=np.float64)
    voxel = np.asarray(voxel, dtype=np.float64)
    ap = np.asarray(action_potential, dtype=np.float64)
    c = np.asarray(c_voxel, dtype=np.float64)
    l, lx, ly, lz = distances(electrode_xyz, voxel)
    w = c[:, None] / l ** 3
    # (D grad u) . l = sum_k grad_k u * (sum_j l_j D_jk)
    Dm = np.asarray(D, dtype=np.float64).reshape(3, 3, -1)
    lv = (lx, ly, lz)
    Wk = [w * sum(lv[j] * Dm[j, k][:, None] for j in range(3)) for k in range(3)]
    T = ap.shape[0]
    out = np.zeros((T, electrode_xyz.shape[0]))
    for t0 in range(0, T, t_chunk):
        g = gradients(ap[t0:t0 + t_chunk], neighbor_id_2d, Delta)
        out[t0:t0 + t_chunk] = g[0] @ Wk[0] + g[1] @ Wk[1] + g[2] @ Wk[2]
    return out


def normalize(egm, typical_magnitude=1.0):
    voltage = np.ptp(egm, axis=0)
    return egm * (typical_magnitude / np.mean(voltage))


def electrogram(electrode_xyz, voxel, D, c_voxel, action_potential, Delta, neighbor_id_2d):
    return normalize(electrogram_raw(electrode_xyz, voxel, D, c_voxel, action_potential, Delta,
                                     neighbor_id_2d))


def electrogram_bruteforce(electrode_xyz, voxel, D, c_voxel, action_potential, Delta,
                           neighbor_id_2d):
    """Triple loop over (t, e, v) of legacy/compute_unipolar_electrogram_cpu.py:22-32."""
    T, N = action_potential.shape
    E = electrode_xyz.shape[0]
    nb = neighbor_id_2d
    out = np.zeros((T, E))
    for t in range(T):
        for v in range(N):
            g = [0.0, 0.0, 0.0]
            for axis in range(3):
                p, m = nb[v, 2 * axis], nb[v, 2 * axis + 1]
                if p >= 0 and m >= 0:
                    g[axis] = (action_potential[t, p] - action_potential[t, m]) / (2 * Delta)
            Dv = np.asarray(D)[:, v].reshape(3, 3)
            Dg = Dv @ np.asarray(g)
            for e in range(E):
                lvec = voxel[v] - electrode_xyz[e]
                l = max(float(np.sqrt(lvec @ lvec)), 1.0)
                out[t, e] += c_voxel[v] / l ** 3 * float(Dg @ lvec)
    return out


def electrode_lat(electrode_signal):
    """Electrode activation time of utility/lat_map.py:20-27: per electrode (axis 1) the time
    index of the steepest downstroke, argmax_t -(s[t+1] - s[t]) (first index on ties, as
    NumPy's argmax), shifted so that the earliest electrode is 0."""
    s = np.asarray(electrode_signal)
    steep = -np.diff(s, axis=0)
    idx = np.argmax(steep, axis=0)
    return idx - np.min(idx)
