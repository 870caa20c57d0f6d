"""The 2D Aliev-Panfilov sheet workload against its committed golden fixture.

tests/golden/sheet_ap.npz is written by tests/golden/make_sheet_fixture.py with the FP64
oracle.  CPU: the fixture is reproduced and obeys the mirror symmetry of the centred
stimulus.  GPU (through the C ABI): the plain FP64 form matches it to 1e-9 relative in u
at the probes and exactly in activation steps; the bricked FP32 form is within 1 step.
"""
import os

import numpy as np
import pytest

from pyheartsim_amd import configs

FIX = os.path.join(os.path.dirname(__file__), "golden", "sheet_ap.npz")


@pytest.fixture(scope="module")
def golden():
    return dict(np.load(FIX))


@pytest.fixture(scope="module")
def work():
    w = configs.sheet_ap()
    w["P_2d"] = configs.build_P_2d(w["geometry"], w["sim"], w["hm"])
    return w


def test_fixture_symmetry(golden):
    lat = golden["lat0"].astype(np.int64).reshape(128, 128)     # x outer, y inner
    assert np.all(lat >= 0)
    # square sheet and patch: x <-> y symmetric everywhere
    assert np.array_equal(lat, lat.T)
    # the patch is centred on index 64 while the edges sit at 0 and 127, so the mirror
    # i -> 128 - i holds away from the edges (before any boundary is felt)
    i = np.arange(24, 105)
    assert np.array_equal(lat[np.ix_(i, i)], lat[np.ix_(128 - i, i)])
    # the activation front moves outward: LAT grows with distance from the centre
    assert np.all(np.diff(lat[64, 64:]) >= 0) and np.all(np.diff(lat[64, :65]) <= 0)


def test_oracle_reproduces_fixture(golden, work):
    from oracle import sim_oracle
    P = work["P_2d"]
    r = sim_oracle.simulate(P.shape[0], P, work["geometry"], work["sim"], work["arr"],
                            probes=golden["probes"], record_samples=False)
    assert np.array_equal(r["probe_u"], golden["probe_u"])
    assert np.array_equal(r["lat"][0].astype(np.int16), golden["lat0"])


@pytest.mark.gpu
def test_gpu_forms_match_fixture(cuda, golden, work):
    from pyheartsim_amd import runtime
    P = work["P_2d"]
    prep = runtime.prepare(P.shape[0], P, work["geometry"], work["sim"], work["arr"])
    plain = runtime.PlainSim(prep, tol=1e-15, probes=golden["probes"])
    plain.run()
    pu = plain.probe_u.cpu().numpy()
    assert np.max(np.abs(pu - golden["probe_u"])) <= 1e-9 * np.max(np.abs(golden["probe_u"]))
    assert np.array_equal(plain.lat.cpu().numpy()[0].astype(np.int16), golden["lat0"])
    assert np.max(np.abs(plain.su.cpu().numpy()[-1] - golden["u_last"])) <= 1e-9
    arena = runtime.Arena(prep.n_voxel, 1, runtime.SolverConfig(brick=(8, 8, 1)))
    slot = arena.add(prep)
    arena.step(prep.n_steps)
    _, _, lat = arena.outputs(slot)
    d = np.abs(lat.cpu().numpy()[0].astype(np.int64) - golden["lat0"].astype(np.int64))
    assert d.max() <= 1
