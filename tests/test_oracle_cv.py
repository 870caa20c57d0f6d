"""Pins of the conduction-velocity oracle (utility/conduction_velocity.py:23-70) by what the
mathematics fixes: exact planar fits, the closed-form minimum-norm solution on a flat sheet,
an independent centred normal-equation solve, the NaN rules and the inclusive radius."""
import numpy as np

from oracle import cv_oracle


def _unit(v):
    v = np.asarray(v, dtype=np.float64)
    return v / np.linalg.norm(v)


def planar_case(n=300, seed=0):
    rng = np.random.default_rng(seed)
    xyz = rng.uniform([-10, -10, -3], [10, 10, 3], size=(n, 3)) + np.array([20.0, -5.0, 12.0])
    nhat = _unit(rng.normal(size=3))
    speed = 0.8  # mm/ms
    lat = 3.0 + xyz @ nhat / speed
    return xyz, lat, nhat, speed


def test_planar_activation_gives_wave_direction_times_speed():
    xyz, lat, nhat, speed = planar_case()
    vec, mag, mean = cv_oracle.conduction_velocity(xyz, lat)
    ok = np.isfinite(mag)
    assert ok.sum() > 250
    np.testing.assert_allclose(vec[ok], np.broadcast_to(nhat * speed, vec[ok].shape), rtol=1e-9,
                               atol=1e-12)
    np.testing.assert_allclose(mag[ok], speed, rtol=1e-9)
    np.testing.assert_allclose(mean, speed, rtol=1e-9)


def flat_sheet(z0=5.0, n=24, step=1.0):
    i, j = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    x = i.ravel() * step - 3.0
    y = j.ravel() * step + 2.0
    xyz = np.column_stack((x, y, np.full(x.shape, z0)))
    a, b, d = 0.7, -0.4, 2.5
    return xyz, a * xyz[:, 0] + b * xyz[:, 1] + d, (a, b, d, z0)


def test_flat_sheet_minimum_norm_solution():
    """Coplanar points z = z0: A = [x, y, z0, 1] has the null vector (0, 0, 1, -z0); lstsq's
    minimum-norm solution splits the intercept d as c = d z0 / (1 + z0^2)."""
    xyz, lat, (a, b, d, z0) = flat_sheet()
    vec, mag, _ = cv_oracle.conduction_velocity(xyz, lat)
    g = np.array([a, b, d * z0 / (1 + z0 * z0)])
    assert np.isfinite(mag).all()
    np.testing.assert_allclose(vec, np.broadcast_to(g / (g @ g), vec.shape), rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(mag, 1 / np.linalg.norm(g), rtol=1e-9)


def test_curved_activation_matches_centred_normal_equations():
    """Point-source activation (not exactly planar): the fitted gradient equals the solution
    of the normal equations on coordinates centred at each neighbourhood's mean."""
    rng = np.random.default_rng(4)
    k = rng.integers(0, 40, size=(600, 3)) * np.array([0.4, 0.4, 0.4])
    xyz = np.unique(k + np.array([10.0, 4.0, -7.0]), axis=0)
    src = np.array([14.0, 9.0, -3.0])
    lat = np.round(np.linalg.norm(xyz - src, axis=1) / 0.6)  # integer LAT, as from the EGM
    vec, mag, _ = cv_oracle.conduction_velocity(xyz, lat)
    r2 = 64.0
    checked = 0
    for i in range(0, len(xyz), 37):
        d = xyz - xyz[i]
        m = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2] <= r2
        if m.sum() < 4:
            continue
        q = xyz[m] - xyz[m].mean(axis=0)
        t = lat[m] - lat[m].mean()
        g = np.linalg.solve(q.T @ q, q.T @ t)
        if g @ g <= 1e-8:
            assert np.isnan(mag[i])
            continue
        np.testing.assert_allclose(vec[i], g / (g @ g), rtol=1e-8, atol=1e-12)
        checked += 1
    assert checked > 5


def test_nan_rules_and_mean():
    # isolated points (more than 8 mm apart): fewer than 4 in range
    xyz = np.array([[0, 0, 0], [20, 0, 0], [0, 20, 0], [0, 0, 20], [40, 40, 40]], dtype=float)
    vec, mag, mean = cv_oracle.conduction_velocity(xyz, np.arange(5.0))
    assert np.isnan(vec).all() and np.isnan(mag).all() and np.isnan(mean)
    # constant activation: zero gradient
    xyz, _, _, _ = planar_case(n=100, seed=2)
    vec, mag, mean = cv_oracle.conduction_velocity(xyz, np.full(100, 7.0))
    assert np.isnan(mag).all() and np.isnan(mean)


def test_radius_is_inclusive_on_lattice_ties():
    xyz = np.array([[0, 0, 0], [8, 0, 0], [0, 8, 0], [0, 0, 8], [-8, 0, 0]], dtype=float)
    lat = np.array([0.0, 5.0, 3.0, 1.0, -4.0])
    vec, mag, _ = cv_oracle.conduction_velocity(xyz, lat)
    assert np.isfinite(mag[0])          # itself plus four points at exactly 8 mm
    assert np.isnan(mag[1:]).all()      # the others are sqrt(128) mm apart
