"""``run_simulation`` of simulation_individual.py:31-92 on the device path.

Same input dictionary (``save_result_flag``, ``result_folder``, ``geometry_data``,
``simulation_parameters``, ``arrhythmia_parameters``, ``heart_model_parameters``, and
``name_prefix`` when saving) and the same steps: fibres and P_2d (equation_parts), the
simulation (``simulation_gpu.compute``), the unipolar electrogram of the electrode voxels
when ``compute_electrogram_flag`` is 1, and the results file
``<result_folder>/<name_prefix>_simulation_results.npz``.  Plotting and the CUDA
availability check are left out: the device path is the only path.  Returns the results
dictionary (the reference returns nothing).
"""
from __future__ import annotations

from pathlib import Path

from . import io
from .simulation import equation_parts, simulation_gpu, unipolar_electrogram_gpu


def run_simulation(input_arguments, **compute_options):
    geometry_data = input_arguments["geometry_data"]
    simulation_parameters = input_arguments["simulation_parameters"]
    arrhythmia_parameters = input_arguments["arrhythmia_parameters"]
    heart_model_parameters = input_arguments["heart_model_parameters"]
    n_voxel = geometry_data["voxel"].shape[0]
    D0 = equation_parts.load_fiber(n_voxel)
    P_2d = equation_parts.heart_model_equation_parts(simulation_parameters, n_voxel, D0,
                                                     geometry_data["neighbor_id_2d"],
                                                     heart_model_parameters)
    action_potential, h, physical_time = simulation_gpu.compute(
        n_voxel, P_2d, geometry_data, simulation_parameters, arrhythmia_parameters,
        **compute_options)
    electrogram_unipolar = None
    if simulation_parameters["compute_electrogram_flag"] == 1:
        voxel = geometry_data["voxel"]
        electrode_xyz = voxel[simulation_parameters["voxel_id_of_simulation_electrode"], :]
        electrogram_unipolar = unipolar_electrogram_gpu.compute(
            electrode_xyz, voxel, D0, heart_model_parameters["c_voxel"], action_potential,
            geometry_data["Delta"], geometry_data["neighbor_id_2d"])
    results = io.simulation_results(
        simulation_parameters, electrogram_unipolar=electrogram_unipolar,
        action_potential=action_potential, h=h, physical_time=physical_time,
        voxel_id_of_simulation_electrode=geometry_data.get("voxel_id_of_simulation_electrode"))
    if input_arguments.get("save_result_flag", 0) == 1:
        folder = Path(input_arguments["result_folder"])
        io.save_simulation_results(folder / f"{input_arguments['name_prefix']}_simulation_results",
                                   results)
    return results
