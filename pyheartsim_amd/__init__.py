"""pyheartsim_amd: MI355X (gfx950) simulation hot path of PyHeartSim.

    from pyheartsim_amd import simulation
    u, h, t = simulation.simulation_gpu.compute(n_voxel, P_2d, geometry_data,
                                                simulation_parameters, arrhythmia_parameters)
    egm = simulation.unipolar_electrogram_gpu.compute(electrode_xyz, voxel, D0, c_voxel,
                                                      u, Delta, neighbor_id_2d)
"""
__version__ = "0.1.0"
