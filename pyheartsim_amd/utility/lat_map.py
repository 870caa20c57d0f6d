"""Drop-in for utility/lat_map.py:compute_electrode_lat on the device.

``compute_electrode_lat(electrode_signal)`` keeps the reference signature
(utility/lat_map.py:20-27): electrode_signal is (T x E), time along axis 0; the
result is the per-electrode index of the steepest downstroke,
argmax_t -(s[t+1] - s[t]) (first index on ties), minus its minimum, as int64
(E,).  It runs as two small kernels (phs_egm_electrode_lat) on the electrogram
where it already is on the device, so no (T x E) host round trip is needed.
NaN differences order as NumPy's argmax orders them: the first NaN wins.

Keyword-only: ``device_out=True`` returns the int32 torch tensor.
"""
from __future__ import annotations

import numpy as np

from .. import _lib


def compute_electrode_lat(electrode_signal, *, device_out: bool = False, stream=None):
    import torch
    lib = _lib.load()
    if isinstance(electrode_signal, torch.Tensor):
        s = electrode_signal.to(device="cuda", dtype=torch.float32).contiguous()
    else:
        s = torch.from_numpy(np.ascontiguousarray(electrode_signal, dtype=np.float32)).cuda()
    if s.dim() != 2 or s.shape[0] < 2:
        raise ValueError("electrode_signal must be (T >= 2) x E")
    T, E = s.shape
    lat = torch.empty(E, dtype=torch.int32, device="cuda")
    _lib.check(lib.phs_egm_electrode_lat(_lib.ptr(s), T, E, _lib.ptr(lat), _lib.stream_ptr(stream)),
               "phs_egm_electrode_lat")
    if device_out:
        return lat
    return lat.cpu().numpy().astype(np.int64)
