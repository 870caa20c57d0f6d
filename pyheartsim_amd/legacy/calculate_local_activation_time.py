"""Drop-in for legacy/calculate_local_activation_time.py on the device.

``execute_on_action_potential(t_start, a_p, v_gate, cycle_length_percentage)`` and
``find_action_potential_marker(s, threshold)`` keep the reference signatures and results
(:18-66, :67-136):

* markers of a signal: the sorted unique boundary points of its runs above the threshold,
  paired in order (an odd last point dropped); per pair the first index of the largest
  forward difference, kept when that difference is > 0; ``[nan]`` when no sample exceeds
  the threshold;
* the mean cycle length CL over all signals (180 when there is none), the window
  ``[t_start, t_start + round(CL * cycle_length_percentage)]``, the first marker inside it per
  signal, shifted so that the earliest is 1 (NaN where there is none).

a_p is (T x n), time along axis 0.  The scan runs on the device (phs_ap_lat, phs_ap_markers),
one thread per signal, over the samples in FP32 (the simulation's output type).
Keyword-only: ``device_out=True`` returns torch tensors.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from .. import _lib


def _device(a):
    import torch
    if isinstance(a, torch.Tensor):
        return a.to(device="cuda", dtype=torch.float32).contiguous()
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).cuda()


def _scan(a_p, threshold, t_start, fraction, stream=None):
    """lat, per-signal flags, markers, cycle lengths and stats of the columns of a_p."""
    import torch
    lib = _lib.load()
    x = _device(a_p)
    if x.dim() != 2 or x.shape[0] < 1 or x.shape[1] < 1:
        raise ValueError("a_p must be (T x n) with T, n >= 1")
    T, n = x.shape
    nbytes = C.c_int64()
    _lib.check(lib.phs_ap_lat_workspace(n, C.byref(nbytes)), "phs_ap_lat_workspace")
    ws = torch.empty(nbytes.value, dtype=torch.uint8, device="cuda")
    lat = torch.empty(n, dtype=torch.float64, device="cuda")
    stats = torch.empty(5, dtype=torch.float64, device="cuda")
    anyv = torch.empty(n, dtype=torch.uint8, device="cuda")
    sp = _lib.stream_ptr(stream)
    _lib.check(lib.phs_ap_lat(_lib.ptr(x), T, n, n, float(threshold), float(t_start),
                              float(fraction), _lib.ptr(ws), _lib.ptr(lat), _lib.ptr(stats),
                              _lib.ptr(anyv), sp), "phs_ap_lat")
    st = stats.cpu().numpy()
    nm, nc = int(st[3]), int(st[4])
    markers = torch.empty(max(nm, 1), dtype=torch.int32, device="cuda")
    cl = torch.empty(max(nc, 1), dtype=torch.float64, device="cuda")
    _lib.check(lib.phs_ap_markers(_lib.ptr(x), T, n, n, float(threshold), _lib.ptr(ws),
                                  _lib.ptr(markers), _lib.ptr(cl), sp), "phs_ap_markers")
    return lat, anyv, markers[:nm], cl[:nc], st


def execute_on_action_potential(t_start, a_p, v_gate, cycle_length_percentage, *,
                                device_out: bool = False, stream=None):
    """(lat, cycle lengths) of legacy/calculate_local_activation_time.py:18-66."""
    lat, _, _, cl, st = _scan(a_p, v_gate, t_start, cycle_length_percentage, stream)
    if device_out:
        return lat, cl
    n = lat.shape[0]
    cl = cl.cpu().numpy()
    # the reference concatenates integer differences, and a float [nan] for every signal
    # without a marker: the result is integer when every signal has one
    if int(st[3]) - int(st[4]) == n:
        cl = cl.astype(np.int64)
    return lat.cpu().numpy(), cl


def find_action_potential_marker(s, threshold):
    """Markers of one signal (legacy/calculate_local_activation_time.py:67-136)."""
    s = np.asarray(s) if not hasattr(s, "dim") else s
    col = s.reshape(-1, 1)
    _, anyv, markers, _, _ = _scan(col, threshold, 0.0, 0.0)
    if not bool(anyv[0].item()):
        return np.array([np.nan])
    return markers.cpu().numpy().astype(int)
