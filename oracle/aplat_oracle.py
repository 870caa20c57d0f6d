"""FP32/FP64 CPU oracle of the action-potential activation markers -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import this module, and only as the checker.

Restates legacy/calculate_local_activation_time.py:
  * markers(s, thr)            :67-136 find_action_potential_marker -- the boundary points
                               of the runs above the threshold (first and last index of each
                               run), sorted and made unique (a one-sample run gives one point,
                               which shifts the pairing of all later points), paired in order
                               with an odd last point dropped; per pair [a, b] the first index
                               of max(diff(s[a:b+1])), dropped when that maximum is <= 0;
                               [nan] when no sample exceeds the threshold;
  * lat_from_markers(...)      :18-66 execute_on_action_potential -- mean cycle length over
                               all signals (180 if none), window [t_start, t_start +
                               round(CL * fraction)], the first marker inside, shifted so that
                               the earliest is 1.

Pinned by tests/test_oracle_aplat.py: hand-built traces with known upstroke indices, the
one-sample-run pairing shift worked out by hand, the deletion of a decaying start, the window
and shift of execute_on_action_potential on hand-computed cycle lengths, and an independent
run-by-run formulation on random traces without one-sample runs.
"""
from __future__ import annotations

import numpy as np


def markers(s, threshold):
    s = np.asarray(s)
    above = np.flatnonzero(s > threshold)
    if above.size == 0:
        return np.array([np.nan])
    gaps = np.flatnonzero(np.diff(above) > 1)
    points = np.unique(np.concatenate(([above[0]], above[gaps], above[gaps + 1], [above[-1]])))
    if points.size % 2:
        points = points[:-1]
    out, keep = [], []
    for a, b in points.reshape(-1, 2):
        d = np.diff(s[a:b + 1])
        out.append(int(np.argmax(d)) + int(a))
        keep.append(not (np.max(d) <= 0))
    return np.asarray(out, dtype=int)[np.asarray(keep, dtype=bool)]


def lat_from_markers(t_start, a_p, v_gate, cycle_length_percentage):
    """(lat, cycle lengths) of the columns of a_p (T x n)."""
    a_p = np.asarray(a_p)
    n = a_p.shape[1]
    per, cls = [], []
    for i in range(n):
        m = markers(a_p[:, i], v_gate)
        if np.isnan(m).all():
            per.append(np.array([np.nan]))
            cls.append(np.array([np.nan]))
        else:
            per.append(m)
            cls.append(np.diff(m))
    cl = np.concatenate(cls)        # integer unless some signal had no marker
    cl = cl[~np.isnan(cl)]
    CL = np.mean(cl) if cl.size else 180
    lo, hi = t_start, t_start + round(CL * cycle_length_percentage)
    lat = np.full(n, np.nan)
    for i, m in enumerate(per):
        m = m[~np.isnan(m)]
        inside = m[(m >= lo) & (m <= hi)]
        if inside.size:
            lat[i] = inside[0]
    with np.errstate(all="ignore"):
        if np.isfinite(lat).any():
            lat = lat - np.nanmin(lat) + 1
    return lat, cl
