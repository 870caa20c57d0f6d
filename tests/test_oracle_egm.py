"""Pins of the electrogram oracle (oracle/egm_oracle.py).

  * brute-force triple loop over (t, e, v) agrees with the vectorised form;
  * single-dipole closed form on a 3-voxel line: egm = c (g . l_vec) / l^3;
  * one-sided axes give zero gradient; l is clamped to >= 1 mm;
  * the electrogram is linear in u; normalisation makes the mean peak-to-peak 1;
  * the drop-in host helpers agree with the oracle's helpers.
"""
import numpy as np

from oracle import egm_oracle as eo
from pyheartsim_amd import geometry as geo
from pyheartsim_amd.simulation import unipolar_electrogram_equation_parts as ep


def _iso(n):
    D = np.zeros((9, n)); D[0] = D[4] = D[8] = 1.0
    return D


def test_bruteforce_matches_vectorised():
    g = geo.shell((3.0, 2.5, 2.0), 1.0, 0.5, n_holes=1, hole_radius=1.0, seed=4)
    n = g["voxel"].shape[0]
    rng = np.random.default_rng(0)
    ap = rng.uniform(size=(3, n))
    el = rng.uniform(-3, 3, size=(4, 3))
    c = rng.uniform(0.5, 2, size=n)
    D = rng.normal(size=(9, n))
    a = eo.electrogram_raw(el, g["voxel"], D, c, ap, g["Delta"], g["neighbor_id_2d"])
    b = eo.electrogram_bruteforce(el, g["voxel"], D, c, ap, g["Delta"], g["neighbor_id_2d"])
    assert np.max(np.abs(a - b)) <= 1e-12 * np.max(np.abs(b))


def test_single_dipole_closed_form_and_one_sided_zero():
    d = 0.4
    ijk = np.array([[0, 0, 0], [1, 0, 0], [2, 0, 0]])
    g = geo._geometry(ijk, d)
    ap = np.array([[0.0, 0.3, 1.0]])
    el = np.array([[0.4, 5.0, 2.0]])
    c = np.array([2.0, 3.0, 5.0])
    out = eo.electrogram_raw(el, g["voxel"], _iso(3), c, ap, d, g["neighbor_id_2d"])
    lvec = g["voxel"][1] - el[0]
    l = np.linalg.norm(lvec)
    grad_x = (1.0 - 0.0) / (2 * d)        # only voxel 1 has both x neighbours
    assert np.isclose(out[0, 0], 3.0 * grad_x * lvec[0] / l ** 3, rtol=1e-14)
    gx, gy, gz = eo.gradients(ap, g["neighbor_id_2d"], d)
    assert gx[0, 0] == 0 and gx[0, 2] == 0 and np.all(gy == 0) and np.all(gz == 0)


def test_clamp_linearity_normalisation():
    g = geo.shell((3.0, 2.5, 2.0), 1.0, 0.5, n_holes=0, seed=5)
    n = g["voxel"].shape[0]
    rng = np.random.default_rng(1)
    el = g["voxel"][:2].copy()            # electrodes on voxels: l clamped to 1 mm
    l, lx, ly, lz = eo.distances(el, g["voxel"])
    assert l.min() == 1.0 and lx[0, 0] == 0.0
    u1, u2 = rng.uniform(size=(4, n)), rng.uniform(size=(4, n))
    c = np.ones(n)
    f = lambda u: eo.electrogram_raw(el, g["voxel"], _iso(n), c, u, 0.5, g["neighbor_id_2d"])
    assert np.allclose(f(2 * u1 - 3 * u2), 2 * f(u1) - 3 * f(u2), rtol=1e-12, atol=1e-12)
    e = eo.normalize(f(u1))
    assert np.isclose(np.mean(np.ptp(e, axis=0)), 1.0, rtol=1e-14)


def test_dropin_helpers_agree_with_oracle():
    g = geo.shell((3.0, 2.5, 2.0), 1.0, 0.5, n_holes=1, hole_radius=1.0, seed=6)
    n = g["voxel"].shape[0]
    rng = np.random.default_rng(2)
    el = rng.uniform(-3, 3, size=(5, 3))
    for a, b in zip(ep.compute_electrode_distances(el, g["voxel"]), eo.distances(el, g["voxel"])):
        assert np.array_equal(a, b)
    ap = rng.uniform(size=(6, n))
    for a, b in zip(ep.compute_gradients(ap, g["neighbor_id_2d"], 0.5),
                    eo.gradients(ap, g["neighbor_id_2d"], 0.5)):
        assert np.array_equal(a, b)
    x = rng.normal(size=(7, 5))
    assert np.allclose(ep.normalize_electrogram_magnitude(x), eo.normalize(x))


def test_electrode_lat_known_answers():
    # electrode 0: one steep fall between t=5 and t=6; electrode 1: two equal falls at t=2 and
    # t=7 (first index wins); electrode 2: a ramp down whose steepest step is t=8 -> t=9
    T = 12
    s = np.zeros((T, 3))
    s[:6, 0] = 1.0
    s[:3, 1] = 2.0
    s[3:8, 1] = 1.0
    s[:, 2] = -np.arange(T) * 0.1
    s[9:, 2] -= 0.5
    lat = eo.electrode_lat(s)
    assert lat.tolist() == [5 - 2, 2 - 2, 8 - 2]
    # a single electrode is always 0; constant signals pick t = 0
    assert eo.electrode_lat(s[:, :1]).tolist() == [0]
    assert eo.electrode_lat(np.ones((5, 4))).tolist() == [0, 0, 0, 0]
