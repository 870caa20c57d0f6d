"""Drop-in for utility/conduction_velocity.py:compute on the device.

``compute(simulation_results, geometry_data)`` keeps the reference signature and returns
(utility/conduction_velocity.py:23-70):

* ``conduction_velocity_vectors`` (n x 3) float64,
* ``conduction_velocity_magnitudes`` (n,) float64,
* ``conduction_velocity_mean`` float64 (mean over the non-NaN magnitudes).

The points are the electrode voxels ``geometry_data['voxel'][voxel_id_of_simulation_electrode]``
and their activation times come from the electrogram (``lat_map.compute_electrode_lat``).
For each point, the plane lat = a x + b y + c z + d is fitted to the points within 8 mm
(minimum-norm least squares, as np.linalg.lstsq); v = g / |g|^2 with g = (a, b, c), and
|v| = 1 / |g|.  Fewer than 4 points in range, or |g|^2 <= 1e-8, give NaN.  The electrogram
can stay on the device: electrode LAT and the fits run there (phs_egm_electrode_lat,
phs_conduction_velocity), and only the n-point results come back.

Keyword-only: ``radius`` (mm, default 8.0 as in the reference), ``device_out=True`` returns
torch tensors (vectors, magnitudes, mean as a 1-element tensor).
"""
from __future__ import annotations

import numpy as np

from .. import _lib
from . import lat_map

NEIGHBOR_RADIUS_MM = 8.0  # utility/conduction_velocity.py:37


def from_points(xyz, lat, *, radius: float = NEIGHBOR_RADIUS_MM, device_out: bool = False,
                stream=None):
    """Conduction velocity at points `xyz` (n x 3, mm) with activation times `lat` (n)."""
    import torch
    lib = _lib.load()

    def dev(a):
        if isinstance(a, torch.Tensor):
            return a.to(device="cuda", dtype=torch.float64).contiguous()
        return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()

    p = dev(xyz)
    t = dev(lat)
    if p.dim() != 2 or p.shape[1] != 3 or t.dim() != 1 or t.shape[0] != p.shape[0]:
        raise ValueError("xyz must be (n, 3) and lat (n,)")
    n = p.shape[0]
    if n == 0:
        raise ValueError("no points")
    vec = torch.empty((n, 3), dtype=torch.float64, device="cuda")
    mag = torch.empty(n, dtype=torch.float64, device="cuda")
    mean = torch.empty(1, dtype=torch.float64, device="cuda")
    _lib.check(lib.phs_conduction_velocity(_lib.ptr(p), _lib.ptr(t), n, float(radius),
                                           _lib.ptr(vec), _lib.ptr(mag), _lib.ptr(mean),
                                           _lib.stream_ptr(stream)), "phs_conduction_velocity")
    if device_out:
        return vec, mag, mean
    return vec.cpu().numpy(), mag.cpu().numpy(), np.float64(mean.item())


def compute(simulation_results, geometry_data, *, radius: float = NEIGHBOR_RADIUS_MM,
            device_out: bool = False, stream=None):
    """utility/conduction_velocity.py:23-70 with the fits on the device."""
    lat = lat_map.compute_electrode_lat(simulation_results["electrogram_unipolar"],
                                        device_out=True, stream=stream)
    ids = np.asarray(geometry_data["voxel_id_of_simulation_electrode"])
    xyz = np.asarray(geometry_data["voxel"])[ids, :]
    return from_points(xyz, lat, radius=radius, device_out=device_out, stream=stream)
