"""Reference-style parameter dicts and the seeded workloads.

``assign_simulation_parameters`` mirrors configuration.py:59-159: the focal
defaults (pacing start 10 ms, cycle length 300 ms), the Mitchell-Schaeffer
and Aliev-Panfilov parameter sets, c = 4.0 (MS) / 1.6 (AP), and the AP time
scaling by 6 (dt, t_final, pacing times divided; c multiplied).  Keyword
overrides replace a default before the time scaling, exactly where the
reference would read it.

Workloads (all seeded):
  sheet_ap   2D Aliev-Panfilov sheet 128 x 128 at 1 mm, 2000 steps of
             dt = 0.5/6, central 5 x 5 stimulus for 2 model-time units.
  shell_ms   Mitchell-Schaeffer ellipsoidal shell (25/20/18 mm, 3 mm wall,
             4 vein holes of 6 mm) at 0.4 mm, dt 0.05 ms, tau_close 150,
             focal pacing CL 500 ms for 2000 ms, 64 basket electrodes.
             The S1 site is the set of voxels within ``pacing_radius`` (2 mm)
             of a seeded outer-surface voxel: a single 0.4 mm voxel driven
             with J = 1 is drained by diffusion and never excites (the
             reference carries a disabled neighbour expansion of the S1 site
             for the same purpose, simulation/pacing.py:110-119).
  shell_fine the shell_ms physics at 0.2 mm with 3 scar blobs.
  ragged     64 shells of uneven size for one batched launch.
"""
from __future__ import annotations

import numpy as np

from . import geometry as geo
from .simulation.equation_parts import heart_model_equation_parts, load_fiber

MS_DEFAULT = dict(tau_in=0.3, tau_out=6, tau_open=120, tau_close=80, v_gate=0.13)
AP_DEFAULT = dict(k=8.0, a=0.15, epsilon_0=0.002, mu1=0.2, mu2=0.3)


def assign_simulation_parameters(geometry_data, s1, s2=None, node_flag=None, *,
                                 heart_model_flag=0, arrhythmia_flag=0, t_final=1000.0,
                                 dt=0.5, pacing_start_time=10.0, pacing_cycle_length=300.0,
                                 s1_t=0.0, s1_s2_delta_t=0.0, ms=None, ap=None, c=None,
                                 geometry_flag=1):
    n_voxel = geometry_data["voxel"].shape[0]
    if node_flag is None:
        node_flag = np.zeros(n_voxel, dtype=np.int64)
    sim = {
        "geometry_flag": geometry_flag,
        "compute_electrogram_flag": 1,
        "save_action_potential_of_all_voxel_flag": 0,
        "t_final": t_final,
        "dt": dt,
        "heart_model_flag": heart_model_flag,
        "arrhythmia_flag": arrhythmia_flag,
    }
    arr = {
        "pacing_start_time": pacing_start_time,
        "pacing_cycle_length": pacing_cycle_length,
        "s1_pacing_voxel_id": s1,
        "s2_pacing_voxel_id": s2 if s2 is not None else [],
        "s1_t": s1_t,
        "s1_s2_delta_t": s1_s2_delta_t,
    }
    msp = dict(MS_DEFAULT, **(ms or {}))
    app = dict(AP_DEFAULT, **(ap or {}))
    ones = np.ones(n_voxel)
    if heart_model_flag == 0:
        hm = {
            "tau_in_voxel": ones * msp["tau_in"],
            "tau_out_voxel": ones * msp["tau_out"],
            "tau_open_voxel": ones * msp["tau_open"],
            "tau_close_voxel": ones * msp["tau_close"],
            "v_gate_voxel": ones * msp["v_gate"],
            "c_voxel": ones * (4.0 if c is None else c),
        }
        ts = 1
    else:
        hm = {
            "k_voxel": ones * app["k"],
            "a_voxel": ones * app["a"],
            "epsilon_0_voxel": ones * app["epsilon_0"],
            "mu1_voxel": ones * app["mu1"],
            "mu2_voxel": ones * app["mu2"],
            "c_voxel": ones * (1.6 if c is None else c),
        }
        ts = 6
    sim["time_scale"] = ts
    sim["t_final"] = sim["t_final"] / ts
    sim["dt"] = sim["dt"] / ts
    for k in ("pacing_start_time", "pacing_cycle_length", "s1_t", "s1_s2_delta_t"):
        arr[k] = arr[k] / ts
    arr["node_flag"] = node_flag
    hm["c_voxel"] = hm["c_voxel"] * ts
    return sim, arr, hm


def build_P_2d(geometry_data, sim, hm):
    n = geometry_data["voxel"].shape[0]
    return heart_model_equation_parts(sim, n, load_fiber(n), geometry_data["neighbor_id_2d"], hm)


def sheet_ap(n_steps: int = 2000, nx: int = 128, ny: int = 128, n_probes: int = 16,
             probe_seed: int = 1):
    """2D Aliev-Panfilov sheet.  The reference's current-injection pacing
    (J = 1, pacing.py:127-132) is applied to the central 5 x 5 patch from
    t = 0 for 2 model-time units (flag 4 single S1 at s1_t = 0)."""
    g = geo.sheet(nx, ny, 1.0)
    cx, cy = nx // 2, ny // 2
    ijk = g["ijk"]
    patch = np.nonzero((np.abs(ijk[:, 0] - cx) <= 2) & (np.abs(ijk[:, 1] - cy) <= 2))[0]
    dt = 0.5  # physical ms before the AP time scaling (configuration.py:75)
    sim, arr, hm = assign_simulation_parameters(
        g, patch, None, None, heart_model_flag=1, arrhythmia_flag=4,
        t_final=n_steps * dt, dt=dt, pacing_start_time=0.0, pacing_cycle_length=0.0,
        s1_t=0.0, geometry_flag=0)
    sim["pacing_duration_override"] = 2.0
    probes = np.sort(np.random.default_rng(probe_seed).choice(g["voxel"].shape[0], n_probes,
                                                                replace=False))
    return dict(name="sheet_ap", geometry=g, sim=sim, arr=arr, hm=hm, probes=probes,
                n_steps=n_steps)


AP_SHELL_STIMULUS = 3.0  # model-time units of the Aliev-Panfilov shells' stimulus


def shell_ms(delta: float = 0.4, t_final: float = 2000.0, seed: int = 0, n_electrodes: int = 64,
             cycle_length: float = 500.0, dt: float = 0.05, tau_close: float = 150.0,
             semi_axes=(25.0, 20.0, 18.0), wall: float = 3.0, n_scar: int = 0,
             scar_radius=(3.0, 6.0), zero_c_in_scar: bool = False, heart_model_flag: int = 0,
             geometry=None, pacing_radius: float = 2.0):
    g = geometry if geometry is not None else geo.shell(semi_axes, wall, delta, 4, 6.0, seed)
    centre = geo.pick_pacing_site(g, seed + 1)
    s1 = geo.pacing_region(g, centre, pacing_radius)
    node_flag = np.zeros(g["voxel"].shape[0], dtype=np.int64)
    if n_scar:
        node_flag = geo.add_scar(g, n_scar, scar_radius, seed + 2, node_flag)
        node_flag[s1] = 0
    sim, arr, hm = assign_simulation_parameters(
        g, s1, None, node_flag, heart_model_flag=heart_model_flag, arrhythmia_flag=0,
        t_final=t_final, dt=dt,   # physical ms; the AP time scaling divides both by 6
        pacing_start_time=10.0, pacing_cycle_length=cycle_length,
        ms=dict(tau_close=tau_close))
    if zero_c_in_scar:
        hm["c_voxel"] = np.where(node_flag == 4, 0.0, hm["c_voxel"])
    if heart_model_flag == 1:
        # Aliev-Panfilov: the reference's 5 / time_scale stimulus (pacing.py:127-132) on the
        # 2 mm pacing site is drained by diffusion (c = 9.6) before the upstroke and never
        # excites the shell; 3 model-time units (18 ms) do, as the 2D sheet's 2 units do
        sim["pacing_duration_override"] = AP_SHELL_STIMULUS
    electrodes = geo.fibonacci_electrodes(n_electrodes, 15.0)
    return dict(name="shell_ms", geometry=g, sim=sim, arr=arr, hm=hm, electrodes=electrodes)


def shell_fine(delta: float = 0.2, t_final: float = 2000.0, seed: int = 0,
               scar_fraction: float = 0.15):
    """shell_ms physics at 0.2 mm with 3 seeded scar blobs (radii drawn in 3-6 mm, then
    scaled together to cover `scar_fraction` of the wall), zero conductivity in the scar for
    both the diffusion (node_flag 4) and the electrogram (c_voxel 0), 256 electrodes."""
    w = shell_ms(delta=delta, t_final=t_final, seed=seed, n_electrodes=256)
    g = w["geometry"]
    nf = geo.add_scar_fraction(g, 3, (3.0, 6.0), scar_fraction, seed + 2)
    nf[np.atleast_1d(w["arr"]["s1_pacing_voxel_id"])] = 0
    w["arr"]["node_flag"] = nf
    w["hm"]["c_voxel"] = np.where(nf == 4, 0.0, w["hm"]["c_voxel"])
    return w


def ragged_params(n_sims: int = 64, t_final: float = 1000.0, seed: int = 7, delta: float = 0.4,
                  axis_range=(15.0, 30.0), wall_range=(2.0, 4.0)):
    """Keyword arguments of `shell_ms` for each member of the batched workload: semi-axes
    U(15, 30) mm (sorted, largest first), wall U(2, 4) mm, cycle length U(250, 600) ms, 0-3
    scar blobs and a seeded pacing site, all drawn from one generator in member order."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n_sims):
        ax = np.sort(rng.uniform(*axis_range, size=3))[::-1]
        wall = rng.uniform(*wall_range)
        cl = rng.uniform(250.0, 600.0)
        n_scar = int(rng.integers(0, 4))
        out.append(dict(delta=delta, t_final=t_final, seed=int(rng.integers(1 << 30)),
                        semi_axes=tuple(float(a) for a in ax), wall=float(wall),
                        cycle_length=float(cl), n_scar=n_scar))
    return out


def ragged(n_sims: int = 64, t_final: float = 1000.0, seed: int = 7, delta: float = 0.4,
           axis_range=(15.0, 30.0), wall_range=(2.0, 4.0)):
    """The 64 uneven shells of the batched workload (see `ragged_params`)."""
    return [shell_ms(**kw) for kw in ragged_params(n_sims, t_final, seed, delta, axis_range,
                                                   wall_range)]
