"""CPU tests of the host side: geometry, P_2d packing, input preparation."""
import numpy as np
import pytest

from pyheartsim_amd import _lib, configs, geometry as geo, runtime
from pyheartsim_amd.simulation import equation_parts as eqp
from tests.helpers import small_shell


def test_sheet_matches_reference_ordering():
    g = geo.sheet(128, 128, 1.0)
    v = g["voxel"]
    # tool/create_geometry.py:47-59: x from -64..63 outer, y inner, z = 0
    assert v.shape == (16384, 3)
    assert np.array_equal(v[0], [-64, -64, 0]) and np.array_equal(v[1], [-64, -63, 0])
    assert np.array_equal(v[128], [-63, -64, 0])
    nb = g["neighbor_id_2d"]
    assert nb[0, 0] == 128 and nb[0, 1] == -1 and nb[0, 2] == 1 and nb[0, 3] == -1
    assert np.all(nb[:, 4:6] == -1)
    assert nb[129, 6] == 129 + 128 + 1     # +x+y


def test_neighbour_table_symmetric_and_lattice_exact():
    g = geo.shell((6.0, 5.0, 4.0), 1.5, 0.4, n_holes=2, hole_radius=2.0, seed=3)
    nb = g["neighbor_id_2d"]
    ijk = g["ijk"]
    for k in range(18):
        m = nb[:, k] >= 0
        d = ijk[nb[m, k]] - ijk[m]
        assert np.all(d == geo.NEIGHBOR_OFFSETS[k])
    for k in range(0, 6, 2):
        m = nb[:, k] >= 0
        assert np.array_equal(nb[nb[m, k], k + 1], np.nonzero(m)[0])


def test_P2d_isotropic_closed_form():
    w = small_shell()
    P = w["P_2d"]
    nb = w["geometry"]["neighbor_id_2d"]
    assert np.array_equal(P[:, :6], 4.0 * (nb[:, :6] >= 0))
    assert np.all(P[:, 6:15] == 0)
    assert np.all(P[:, 15:20] == [120, 150, 0.3, 6, 0.13])
    assert np.all(P[:, 20] == 4.0)
    D0 = eqp.load_fiber(5)
    assert len(D0) == 5 and np.array_equal(D0[3], np.eye(3))
    # explicit list of tensors gives the same packing
    n = P.shape[0]
    P2 = eqp.heart_model_equation_parts(w["sim"], n, [np.eye(3)] * n, nb, w["hm"])
    assert np.array_equal(P, P2)


def test_prepare_flags_schedule_sampling():
    w = small_shell(n_scar=2, temp_block=True)
    P = w["P_2d"]
    prep = runtime.prepare(P.shape[0], P, w["geometry"], w["sim"], w["arr"])
    nf = w["arr"]["node_flag"]
    assert np.array_equal((prep.vflags & _lib.PHS_VF_SCAR) > 0, nf == 4)
    assert np.array_equal((prep.vflags & _lib.PHS_VF_TEMP_BLOCK) > 0, nf == 3)
    s1 = np.zeros(P.shape[0], bool); s1[w["arr"]["s1_pacing_voxel_id"]] = True
    assert np.array_equal((prep.vflags & _lib.PHS_VF_S1) > 0, s1)
    assert prep.c == 4.0 and prep.c_voxel is None
    assert prep.n_steps == 800 and prep.steps_per_sample == 20 and prep.n_samples == 40
    t = prep.physical_time()
    assert len(t) == prep.n_saved() == 40 and t[1] == np.float32(1.0)
    assert np.array_equal(prep.ijk, w["geometry"]["ijk"] - w["geometry"]["ijk"].min(axis=0)) or \
        np.array_equal(prep.ijk, w["geometry"]["ijk"])


def test_prepare_rejects_anisotropy_keeps_hetero_params():
    w = small_shell()
    P = w["P_2d"].copy()
    P[3, 7] = 0.5
    with pytest.raises(_lib.PhsError, match="PHS_ERR_UNSUPPORTED") as e:
        runtime.prepare(P.shape[0], P, w["geometry"], w["sim"], w["arr"])
    assert e.value.code == _lib.PHS_ERR_UNSUPPORTED
    P = w["P_2d"].copy()
    P[3, 17] = 0.31
    prep = runtime.prepare(P.shape[0], P, w["geometry"], w["sim"], w["arr"])
    assert prep.params_voxel.shape == (5, P.shape[0])
    assert np.array_equal(prep.params_voxel, P[:, 15:20].T)
    assert prep.desc().params_voxel == prep.params_voxel.ctypes.data
    uniform = runtime.prepare(P.shape[0], w["P_2d"], w["geometry"], w["sim"], w["arr"])
    assert uniform.params_voxel is None and uniform.desc().params_voxel is None


def test_lattice_recovered_from_neighbours():
    g = geo.shell((4.0, 3.0, 3.0), 1.0, 0.5, n_holes=1, hole_radius=1.0, seed=2)
    ijk = geo.lattice_from_neighbors(g["neighbor_id_2d"])
    nb2 = geo.neighbor_table(ijk)
    assert np.array_equal(nb2[:, :6], g["neighbor_id_2d"][:, :6])


def test_ap_time_scaling_like_configuration():
    g = geo.sheet(8, 8)
    sim, arr, hm = configs.assign_simulation_parameters(g, 0, heart_model_flag=1)
    assert sim["time_scale"] == 6 and sim["dt"] == 0.5 / 6 and sim["t_final"] == 1000 / 6
    assert arr["pacing_start_time"] == 10 / 6 and np.all(hm["c_voxel"] == 1.6 * 6)


def test_brick_plan_accepts_cut_links_rejects_foreign_neighbours():
    """A table may omit lattice neighbours (asymmetric cuts included); an entry that is not
    the adjacent lattice voxel is refused as unsupported."""
    g = geo.shell((5.0, 4.0, 4.0), 1.0, 0.5, n_holes=1, hole_radius=1.5, seed=2)
    nb = np.ascontiguousarray(g["neighbor_id_2d"][:, :6], dtype=np.int32)
    cut = nb.copy()
    rng = np.random.default_rng(0)
    rows = rng.choice(len(nb), 40, replace=False)
    cut[rows, 0] = -1                       # +x links cut one way only
    perm, masks = runtime.brick_plan(g["ijk"], cut)
    assert sorted(perm.tolist()) == list(range(len(nb)))
    bad = nb.copy()
    v = int(np.nonzero(nb[:, 0] >= 0)[0][0])
    bad[v, 0] = (nb[v, 0] + 7) % len(nb)    # some other voxel
    with pytest.raises(_lib.PhsError, match="not the adjacent lattice voxel"):
        runtime.brick_plan(g["ijk"], bad)
