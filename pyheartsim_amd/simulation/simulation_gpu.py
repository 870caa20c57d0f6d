"""Drop-in for simulation/simulation_gpu.py.

``compute`` keeps the reference signature and return values
(simulation/simulation_gpu.py:261, :409):

    sim_u_voxel, sim_h_voxel, physical_time = compute(
        n_voxel, P_2d, geometry_data, simulation_parameters, arrhythmia_parameters)

with float32 (samples x N) arrays in the caller's voxel order.  Keyword-only
extras select the device form:

    precision="fp32"   bricked FP32 arena (default; reaction and diffusion in
                       FP32, CN right-hand side in FP64)
    precision="fp64"   plain FP64 form (reference algorithm, tight CN solve)
    sweeps, guess, refine, refine_sweeps, brick   solver settings of the FP32 form
                       (defaults: sweeps=0, the automatic count of Chebyshev sweeps;
                       guess=2, the CN correction warm-started from the linear
                       extrapolation of the previous two steps' corrections.  The
                       automatic count is the fewest sweeps that keep the step stable for
                       the geometry's spectral bound, keep the solve error at the stiffest
                       mode under a quarter of CN's damping, damp the stiff modes, and
                       meet an accuracy target that tightens for coarse time steps --
                       6 on the 0.4 mm shell and 17 at 0.2 mm at dt = 0.05 ms, 16 at
                       1 mm and dt = 0.5 ms, each validated at LAT error <= 1 dt against
                       the FP64 oracle or the plain FP64 form)
    return_lat=True    also return the per-voxel activation-time table
                       (max_beats x N step indices, -1 = none)
    device_out=True    return torch device tensors instead of NumPy arrays

Both forms run on the GPU through libphs.so; there is no CPU fallback.  A run whose
transmembrane potential becomes NaN or inf emits a RuntimeWarning naming the first such step
(the device records it; the reference only prints solver non-convergence,
simulation/simulation_gpu.py:250-255).
"""
from __future__ import annotations

import warnings

import numpy as np

from .. import runtime


def compute(n_voxel, P_2d, geometry_data, simulation_parameters, arrhythmia_parameters, *,
            precision: str = "fp32", sweeps: int = 0, guess: int = 2, refine: int = 0,
            refine_sweeps: int = 8, brick=(4, 4, 4), plain_tol: float = 1e-15,
            return_lat: bool = False, device_out: bool = False):
    prep = runtime.prepare(n_voxel, P_2d, geometry_data, simulation_parameters,
                           arrhythmia_parameters, need_lattice=(precision != "fp64"))
    ptime = prep.physical_time()
    if precision == "fp64":
        sim = runtime.PlainSim(prep, tol=plain_tol, sample_fp64=False)
        sim.run()
        su, sh, lat = sim.su, sim.sh, sim.lat
        bad = sim.nonfinite_step()
    elif precision == "fp32":
        solver = runtime.SolverConfig(sweeps=sweeps, guess=guess, refine=refine,
                                      refine_sweeps=refine_sweeps, brick=tuple(brick))
        arena = runtime.Arena(max_voxels=n_voxel, max_sims=1, solver=solver,
                              max_bricks=n_voxel)
        slot = arena.add(prep, samples=True, sample_h=True)
        while arena.steps_done(slot) < prep.n_steps:
            arena.step(min(500, prep.n_steps - arena.steps_done(slot)), graph=True)
        su, sh, lat = arena.outputs(slot)
        bad = arena.nonfinite_step(slot)
    else:
        raise ValueError(f"Unknown precision: {precision}. Use 'fp32' or 'fp64'")
    if bad >= 0:
        warnings.warn(f"transmembrane potential is not finite from step {bad} "
                      f"(model time {bad * prep.dt:g})", RuntimeWarning)
    T = len(ptime)
    su, sh = su[:T], sh[:T]
    if device_out:
        res = (su, sh, ptime)
    else:
        res = (su.cpu().numpy(), sh.cpu().numpy(), ptime)
    if return_lat:
        res = res + (lat if device_out else lat.cpu().numpy(),)
    return res
