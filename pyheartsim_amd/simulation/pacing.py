"""Pacing schedule (drop-in for simulation/pacing.py).

``assign_pacing_parameters`` and ``apply_pacing`` keep the reference's
signatures and return values (simulation/pacing.py:102-134, :136-170).  The
device path never evaluates them per step: ``pacing_schedule`` evaluates the
same FP64 decision rule once per step on the host and packs the result into a
uint8 table that the kernels index by step number.

Schedule bits (one byte per step):
  bit 0  S1 sites stimulated this step
  bit 1  S2 sites stimulated this step
  bit 2  temporary block (node_flag 3) is non-conductive this step
         (simulation/simulation_gpu.py:364-370: ``model_time < 350 * dt``)
"""
from __future__ import annotations

import numpy as np

SCHED_S1 = 1
SCHED_S2 = 2
SCHED_TEMP_BLOCK = 4


def find_out_s2_pacing_voxel_ids_for_rotor_arrhythmia(s1, geometry_data):
    """S2 site for a rotor set-up (simulation/pacing.py:16-100).

    Candidates are the voxels 12-32 mm from the S1 voxel.  They are grouped into connected
    clusters over the 18-neighbour table, by depth-first search from each not yet visited
    candidate in increasing id order.  The largest cluster (the first found among equals)
    is taken.  If it holds more than 300 voxels, S2 is the first 300 voxels of a
    breadth-first search from its first voxel, neighbours in table column order.
    Otherwise S2 is that single first voxel.  Membership tests use boolean masks instead of
    the reference's list scans; the visiting order is the same.
    """
    voxel = np.asarray(geometry_data["voxel"], dtype=np.float64)
    nb = np.asarray(geometry_data["neighbor_id_2d"])
    n = voxel.shape[0]
    d = np.sqrt(np.sum((voxel - voxel[s1, :]) ** 2, axis=1))
    cand = np.nonzero((d >= 12) & (d <= 12 + 20))[0]
    is_cand = np.zeros(n, dtype=bool)
    is_cand[cand] = True
    visited = np.zeros(n, dtype=bool)
    best = []
    for root in cand:
        if visited[root]:
            continue
        cluster = []
        stack = [int(root)]
        visited[root] = True
        while stack:
            cur = stack.pop()
            cluster.append(cur)
            for m in nb[cur]:
                if m >= 0 and is_cand[m] and not visited[m]:
                    visited[m] = True
                    stack.append(int(m))
        if len(cluster) > len(best):
            best = cluster
    first = best[0]
    n_threshold = 300
    if len(best) <= n_threshold:
        return first
    in_best = np.zeros(n, dtype=bool)
    in_best[best] = True
    seen = np.zeros(n, dtype=bool)
    seen[first] = True
    queue = [first]
    head = 0
    out = []
    while head < len(queue) and len(out) < n_threshold:
        cur = queue[head]
        head += 1
        out.append(cur)
        for m in nb[cur]:
            if m >= 0 and in_best[m] and not seen[m]:
                seen[m] = True
                queue.append(int(m))
    return out


def assign_pacing_parameters(arrhythmia_parameters, arrhythmia_flag, n_voxel, neighbor_id_2d,
                             simulation_parameters):
    """Reference simulation/pacing.py:102-134 (same 7-tuple)."""
    s1_sites = arrhythmia_parameters["s1_pacing_voxel_id"]
    s1_t = arrhythmia_parameters["s1_t"]
    s2_t = s1_t + arrhythmia_parameters["s1_s2_delta_t"]
    s2_sites = []
    if arrhythmia_flag in (1, 2, 3, 4, 5, 6):
        s2_sites = arrhythmia_parameters["s2_pacing_voxel_id"]
    J_stim = np.zeros(n_voxel)
    magnitude = 1
    if simulation_parameters["heart_model_flag"] == 0:
        duration = 5
    else:
        duration = 5 / simulation_parameters["time_scale"]
    # Optional extension key (not in the reference): a stimulus duration in
    # model-time units, used by the 2D sheet workload's 2-unit stimulus.
    duration = simulation_parameters.get("pacing_duration_override", duration)
    return J_stim, s1_sites, s2_sites, s1_t, magnitude, duration, s2_t


def stimulus_windows(arrhythmia_parameters, simulation_parameters, arrhythmia_flag, model_time,
                     s1_t, pacing_duration, s2_t):
    """(S1 on, S2 on) at ``model_time``: the rule of simulation/pacing.py:136-170."""
    s1_on = False
    s2_on = False
    if arrhythmia_flag == 0:
        start = arrhythmia_parameters["pacing_start_time"]
        cl = arrhythmia_parameters["pacing_cycle_length"]
        t = model_time
        while t - start > cl:
            t = t - cl
        if start <= t <= start + pacing_duration:
            s1_on = True
    if arrhythmia_flag in (5, 6):
        ts = simulation_parameters["time_scale"]
        f1 = 0 / ts
        f2 = (300 if arrhythmia_flag == 5 else 50) / ts
        if f1 <= model_time <= f1 + pacing_duration:
            s1_on = True
        if f2 <= model_time <= f2 + pacing_duration:
            s2_on = True
    elif arrhythmia_flag in (1, 2, 3, 4):
        if s1_t <= model_time <= s1_t + pacing_duration:
            s1_on = True
    if arrhythmia_flag in (1, 2, 3) and s2_t <= model_time <= s2_t + pacing_duration:
        s2_on = True
    return s1_on, s2_on


def apply_pacing(arrhythmia_parameters, simulation_parameters, arrhythmia_flag, model_time, J_stim,
                 s1_pacing_voxel_id, s2_pacing_voxel_id, s1_t, J_stim_magnitude, pacing_duration,
                 s2_t, sim_u_voxel, sim_h_voxel, neighbor_id_2d):
    """Reference simulation/pacing.py:136-170 (same signature, fills J_stim)."""
    s1_on, s2_on = stimulus_windows(arrhythmia_parameters, simulation_parameters, arrhythmia_flag,
                                    model_time, s1_t, pacing_duration, s2_t)
    if s1_on:
        J_stim[s1_pacing_voxel_id] = J_stim_magnitude
    if s2_on:
        J_stim[s2_pacing_voxel_id] = J_stim_magnitude
    return J_stim


def total_steps(simulation_parameters) -> int:
    """simulation/simulation_gpu.py:355."""
    return int(np.round(float(simulation_parameters["t_final"]) / simulation_parameters["dt"]))


def pacing_schedule(arrhythmia_parameters, simulation_parameters, n_steps: int | None = None,
                    step0: int = 0) -> np.ndarray:
    """uint8 per-step table for steps step0 .. step0+n_steps-1.

    ``model_time = step * dt`` in FP64 exactly as simulation_gpu.py:362, then
    the stimulus rule and the temporary-block window of :364-370.
    """
    dt = simulation_parameters["dt"]
    flag = simulation_parameters["arrhythmia_flag"]
    if n_steps is None:
        n_steps = total_steps(simulation_parameters) - step0
    _, _, _, s1_t, _, duration, s2_t = assign_pacing_parameters(
        arrhythmia_parameters, flag, 0, None, simulation_parameters)
    out = np.zeros(n_steps, dtype=np.uint8)
    for i in range(n_steps):
        model_time = (step0 + i) * dt
        s1_on, s2_on = stimulus_windows(arrhythmia_parameters, simulation_parameters, flag,
                                        model_time, s1_t, duration, s2_t)
        b = (SCHED_S1 if s1_on else 0) | (SCHED_S2 if s2_on else 0)
        if 0 <= model_time < 350 * dt:
            b |= SCHED_TEMP_BLOCK
        out[i] = b
    return out


def site_ids(arrhythmia_parameters, simulation_parameters, n_voxel):
    """S1 and S2 voxel id arrays (int64) as apply_pacing would index them."""
    flag = simulation_parameters["arrhythmia_flag"]
    _, s1, s2, *_ = assign_pacing_parameters(arrhythmia_parameters, flag, n_voxel, None,
                                             simulation_parameters)
    s1 = np.atleast_1d(np.asarray(s1, dtype=np.int64)) if s1 is not None else np.zeros(0, np.int64)
    s2 = np.atleast_1d(np.asarray(s2, dtype=np.int64)) if len(np.atleast_1d(s2)) else np.zeros(0, np.int64)
    return s1, s2
