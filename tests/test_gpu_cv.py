"""GPU parity of the conduction-velocity fit (phs_conduction_velocity, through the C ABI) with
the FP64 oracle of utility/conduction_velocity.py:23-70.

Bar: the same NaN pattern; |v_gpu - v_oracle| <= 1e-8 |v_oracle| per point (both are FP64;
they differ in summation order and in the 4x4 solve: normal equations + Jacobi on the device,
SVD in LAPACK); the mean within 1e-10 relative."""
import numpy as np
import pytest

from oracle import cv_oracle, egm_oracle
from tests.test_oracle_cv import flat_sheet, planar_case

pytestmark = pytest.mark.gpu


def _cases():
    rng = np.random.default_rng(4)
    k = rng.integers(0, 40, size=(600, 3)) * 0.4
    xyz = np.unique(k + np.array([10.0, 4.0, -7.0]), axis=0)
    curved = (xyz, np.round(np.linalg.norm(xyz - np.array([14.0, 9.0, -3.0]), axis=1) / 0.6))
    t = np.linspace(0, 30, 80)
    line = (np.column_stack((t, 2 * t + 1, 0.5 * t - 3)), 0.3 * t + 1)
    fx, fl, _ = flat_sheet(z0=-40.0, n=20, step=0.4)
    sparse = (np.array([[0, 0, 0], [20, 0, 0], [0, 20, 0], [0, 0, 20], [8, 0, 0], [0, 8, 0],
                        [-8, 0, 0], [0, 0, 8]], dtype=float), np.array([0, 1, 2, 3, 4, 5, 6, 7.0]))
    big = rng.uniform(-30, 30, size=(3000, 3))
    big_lat = np.floor(np.linalg.norm(big - np.array([5.0, -3.0, 2.0]), axis=1) / 0.7)
    return {"planar": planar_case()[:2], "flat": flat_sheet()[:2], "curved": curved,
            "line": line, "flat_far": (fx + np.array([50.0, 60.0, 0.0]), fl),
            "sparse_ties": sparse, "constant": (planar_case(n=64)[0], np.full(64, 3.0)),
            "big": (big, big_lat)}


def _check(got, ref):
    v1, m1, mean1 = got
    v0, m0, mean0 = ref
    assert np.array_equal(np.isnan(m1), np.isnan(m0))
    assert np.array_equal(np.isnan(v1), np.isnan(v0))
    ok = np.isfinite(m0)
    if ok.any():
        dv = np.linalg.norm(v1[ok] - v0[ok], axis=1)
        assert np.all(dv <= 1e-8 * np.linalg.norm(v0[ok], axis=1)), float(np.max(dv))
        np.testing.assert_allclose(m1[ok], m0[ok], rtol=1e-8)
        np.testing.assert_allclose(mean1, mean0, rtol=1e-10)
    else:
        assert np.isnan(mean1) and np.isnan(mean0)


@pytest.mark.parametrize("name", ["planar", "flat", "curved", "line", "flat_far", "sparse_ties",
                                  "constant", "big"])
def test_cv_matches_oracle(cuda, name):
    from pyheartsim_amd.utility import conduction_velocity
    xyz, lat = _cases()[name]
    print(f"case {name}: n={len(xyz)}", flush=True)
    _check(conduction_velocity.from_points(xyz, lat), cv_oracle.conduction_velocity(xyz, lat))


def test_cv_dropin_from_electrogram(cuda):
    """compute(simulation_results, geometry_data): electrode LAT from the electrogram on the
    device, then the fits; same as the oracle chain on the host."""
    from pyheartsim_amd.utility import conduction_velocity
    rng = np.random.default_rng(7)
    voxel = np.unique(rng.integers(0, 60, size=(4000, 3)), axis=0) * 0.4
    ids = rng.choice(len(voxel), size=200, replace=False)
    # a downstroke per electrode at a time set by its distance to a source
    T = 400
    onset = np.floor(np.linalg.norm(voxel[ids] - voxel[ids[0]], axis=1) / 0.05).astype(int) % (T - 20)
    t = np.arange(T)[:, None]
    egm = (np.tanh((onset[None, :] - t) * 0.8) + 0.01 * rng.standard_normal((T, len(ids))))
    egm = egm.astype(np.float32)
    results = {"electrogram_unipolar": egm}
    geometry = {"voxel": voxel, "voxel_id_of_simulation_electrode": ids}
    got = conduction_velocity.compute(results, geometry)
    lat = egm_oracle.electrode_lat(egm)
    _check(got, cv_oracle.conduction_velocity(voxel[ids], lat))
