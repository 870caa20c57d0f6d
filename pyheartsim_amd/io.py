"""Reference-compatible files: geometry ``.npz`` in, simulation results ``.npz`` out.

* ``load_geometry`` reads a ``*_mesh.npz`` / ``*_geometry.npz`` as the reference does
  (simulation_individual.py:105-108: every array of the file, ``allow_pickle=False``) and,
  when the file has voxels but no ``neighbor_id_2d``, builds the table with
  ``utility.voxelization.find_neighbor_voxel_ids``.
* ``simulation_results`` / ``save_simulation_results`` produce the dictionary and file of
  simulation_individual.py:77-92 (keys ``geometry_flag``, ``electrogram_unipolar`` or
  ``action_potential_electrode``, and with the save-all flag ``action_potential``, ``h``,
  ``physical_time``), plus ``lat_electrode`` as simulation_individual.py:365 adds it.
"""
from __future__ import annotations

import numpy as np

from .utility.voxelization import find_neighbor_voxel_ids


def load_geometry(path) -> dict:
    with np.load(path, allow_pickle=False) as data:
        g = {k: data[k] for k in data.files}
    if "voxel" in g and "neighbor_id_2d" not in g:
        g["neighbor_id_2d"] = find_neighbor_voxel_ids(g["voxel"])
    return g


def simulation_results(simulation_parameters, *, electrogram_unipolar=None, action_potential=None,
                       h=None, physical_time=None, voxel_id_of_simulation_electrode=None,
                       lat_electrode=None) -> dict:
    res = {"geometry_flag": simulation_parameters["geometry_flag"]}
    if simulation_parameters["compute_electrogram_flag"] == 0:
        res["action_potential_electrode"] = np.asarray(action_potential)[:, voxel_id_of_simulation_electrode]
    elif simulation_parameters["compute_electrogram_flag"] == 1:
        res["electrogram_unipolar"] = electrogram_unipolar
    if simulation_parameters.get("save_action_potential_of_all_voxel_flag", 0) == 1:
        res["action_potential"] = action_potential
        res["h"] = h
        res["physical_time"] = physical_time
    if lat_electrode is not None:
        res["lat_electrode"] = lat_electrode
    return res


def save_simulation_results(path, results: dict):
    np.savez(path, **results)
