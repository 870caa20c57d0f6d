"""Small seeded workloads shared by the CPU and GPU tests."""
import numpy as np

from pyheartsim_amd import configs, geometry as geo


def small_shell(delta=0.5, semi_axes=(9.0, 8.0, 7.0), wall=2.0, t_final=40.0, dt=0.05, seed=3,
                n_scar=0, cycle_length=500.0, heart_model_flag=0, n_electrodes=16, temp_block=False):
    g = geo.shell(semi_axes, wall, delta, n_holes=2, hole_radius=3.0, seed=seed)
    w = configs.shell_ms(delta=delta, t_final=t_final, seed=seed, n_electrodes=n_electrodes,
                         cycle_length=cycle_length, dt=dt, semi_axes=semi_axes, wall=wall,
                         n_scar=n_scar, scar_radius=(1.5, 2.5), heart_model_flag=heart_model_flag,
                         geometry=g)
    if temp_block:
        nf = w["arr"]["node_flag"].copy()
        v = g["voxel"]
        nf[(v[:, 0] > 0) & (np.abs(v[:, 1]) < 1.5)] = 3
        nf[w["arr"]["s1_pacing_voxel_id"]] = 0
        w["arr"]["node_flag"] = nf
    w["P_2d"] = configs.build_P_2d(g, w["sim"], w["hm"])
    w["electrodes"] = geo.fibonacci_electrodes(n_electrodes, 4.0)
    return w


def small_sheet(n=32, n_steps=400):
    w = configs.sheet_ap(n_steps=n_steps, nx=n, ny=n, n_probes=8)
    w["P_2d"] = configs.build_P_2d(w["geometry"], w["sim"], w["hm"])
    return w


def run_oracle(w, **kw):
    from oracle import sim_oracle
    P = w["P_2d"]
    return sim_oracle.simulate(P.shape[0], P, w["geometry"], w["sim"], w["arr"], **kw)
