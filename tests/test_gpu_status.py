"""Divergence report of both device forms (phs_plain_status, phs_arena_slot_status).

A run whose transmembrane potential turns NaN or inf records the first such step on the
device.  Forced from the host with ordinary inputs: a NaN initial value (non-finite from step
0), and a Mitchell-Schaeffer tau_in of 0.01 ms at dt = 0.5 ms, far past the RK4 stability
limit, which overflows after the first stimulus.  A normal run reports -1, and the drop-in
compute warns with the step.
"""
import warnings

import numpy as np
import pytest

from tests.helpers import small_shell

pytestmark = pytest.mark.gpu


def _prep(w):
    from pyheartsim_amd import runtime
    P = w["P_2d"]
    return runtime.prepare(P.shape[0], P, w["geometry"], w["sim"], w["arr"])


def _both(prep):
    from pyheartsim_amd import runtime
    plain = runtime.PlainSim(prep, tol=1e-15, samples=False)
    plain.run()
    arena = runtime.Arena(prep.n_voxel, 1, runtime.SolverConfig())
    slot = arena.add(prep, samples=False)
    arena.step(prep.n_steps)
    return plain.nonfinite_step(), arena.nonfinite_step(slot)


def test_normal_run_reports_none(cuda):
    assert _both(_prep(small_shell(t_final=20.0))) == (-1, -1)


def test_nan_initial_value_reported_at_step_0(cuda):
    prep = _prep(small_shell(t_final=5.0))
    u0 = np.zeros(prep.n_voxel, dtype=np.float32)
    u0[prep.n_voxel // 2] = np.nan
    prep.u0 = u0
    assert _both(prep) == (0, 0)


def test_unstable_reaction_reported(cuda):
    from pyheartsim_amd import configs
    from pyheartsim_amd.simulation import simulation_gpu
    w = small_shell(t_final=60.0, dt=0.5)
    w["hm"]["tau_in_voxel"] = np.full_like(w["hm"]["tau_in_voxel"], 0.01)
    w["P_2d"] = configs.build_P_2d(w["geometry"], w["sim"], w["hm"])
    prep = _prep(w)
    bad_plain, bad_arena = _both(prep)
    first_stim = int(np.argmax(prep.schedule & 1))
    print(f"first stimulus step {first_stim}, non-finite from plain {bad_plain}, "
          f"arena {bad_arena}")
    assert first_stim <= bad_plain < prep.n_steps
    assert first_stim <= bad_arena < prep.n_steps
    P = w["P_2d"]
    with warnings.catch_warnings(record=True) as got:
        warnings.simplefilter("always")
        simulation_gpu.compute(P.shape[0], P, w["geometry"], w["sim"], w["arr"],
                               precision="fp64")
    assert any(issubclass(x.category, RuntimeWarning) and f"step {bad_plain}" in str(x.message)
               for x in got)
