"""Neighbour tables that are not the lattice's own, through the C ABI, against the FP64 oracle.

The reference builds its operator and its gradients from whatever table it is given
(simulation/simulation_gpu.py:146-168, simulation/unipolar_electrogram_equation_parts.py:60-103),
so a table with links missing one way (an asymmetric operator) or with a link to a voxel that
is not adjacent on the lattice is run as given.

Bars: cut links (3 % of the face entries, one way only): plain FP64 within 1e-9 of the oracle
with identical activation times, FP32 arena within 1 step, electrogram within rel-L2 1e-3
(BF16 hi/lo) and 1e-5 (FP32) of the oracle.  A link to a non-adjacent voxel: the plain FP64
form within 1e-9 of the oracle; the bricked form refuses it with PHS_ERR_UNSUPPORTED.
"""
import numpy as np
import pytest

from tests.helpers import run_oracle, small_shell

pytestmark = pytest.mark.gpu


def _with_table(w, nb):
    from pyheartsim_amd import configs
    w = dict(w)
    w["geometry"] = dict(w["geometry"], neighbor_id_2d=nb)
    w["P_2d"] = configs.build_P_2d(w["geometry"], w["sim"], w["hm"])
    return w


def _prep(w):
    from pyheartsim_amd import runtime
    P = w["P_2d"]
    return runtime.prepare(P.shape[0], P, w["geometry"], w["sim"], w["arr"])


def test_cut_links_against_oracle(cuda):
    torch = cuda
    from oracle import egm_oracle
    from pyheartsim_amd import runtime
    from pyheartsim_amd.simulation.unipolar_electrogram_gpu import LeadField
    w0 = small_shell(t_final=40.0, seed=8)
    nb = w0["geometry"]["neighbor_id_2d"].copy()
    rng = np.random.default_rng(4)
    present = np.argwhere(nb[:, :6] >= 0)
    pick = present[rng.choice(len(present), len(present) * 3 // 100, replace=False)]
    nb[pick[:, 0], pick[:, 1]] = -1
    rev = nb[:, :6].copy()
    asym = sum(1 for v, k in pick if w0["geometry"]["neighbor_id_2d"][v, k] >= 0 and
               rev[w0["geometry"]["neighbor_id_2d"][v, k], k ^ 1] >= 0)
    assert asym > 0.9 * len(pick)          # mostly one-way cuts
    w = _with_table(w0, nb)
    ref = run_oracle(w, max_beats=7)
    prep = _prep(w)
    plain = runtime.PlainSim(prep, tol=1e-15)
    plain.run()
    err = np.max(np.abs(plain.su.cpu().numpy() - ref["u"])) / np.max(np.abs(ref["u"]))
    print(f"N={prep.n_voxel} cuts={len(pick)} plain rel err {err:.2e}")
    assert err <= 1e-9
    assert np.array_equal(plain.lat.cpu().numpy().astype(np.int64), ref["lat"][:7])
    arena = runtime.Arena(prep.n_voxel, 1, runtime.SolverConfig())
    slot = arena.add(prep)
    arena.step(prep.n_steps)
    lat = arena.lat(slot).cpu().numpy().astype(np.int64)
    r = ref["lat"][:7]
    assert np.array_equal(lat >= 0, r >= 0)
    m = r >= 0
    d = int(np.max(np.abs(lat[m] - r[m])))
    print(f"fp32 LAT max diff {d} over {int(m.sum())}")
    assert d <= 1
    # the cuts matter: the full table gives other activation times
    assert not np.array_equal(run_oracle(w0, max_beats=7)["lat"][:7], r)
    g = w["geometry"]
    N = prep.n_voxel
    D = np.zeros((9, N))
    D[0] = D[4] = D[8] = 1.0
    el = w["electrodes"]
    e_ref = egm_oracle.electrogram_raw(el, g["voxel"], D, w["hm"]["c_voxel"], ref["u"],
                                       g["Delta"], g["neighbor_id_2d"])
    lf = LeadField(el, g["voxel"], None, w["hm"]["c_voxel"], g["Delta"], g["neighbor_id_2d"])
    s = torch.from_numpy(ref["u"].astype(np.float32)).cuda()
    for prec, bar in (("bf16x3", 1e-3), ("fp32", 1e-5)):
        out = lf.run(s, precision=prec, normalize=False).cpu().numpy()
        rel = float(np.linalg.norm(out - e_ref) / np.linalg.norm(e_ref))
        print(prec, f"egm rel-L2 {rel:.2e}")
        assert rel <= bar


def test_non_adjacent_link(cuda):
    from pyheartsim_amd import _lib, runtime
    w0 = small_shell(t_final=30.0, seed=9)
    nb = w0["geometry"]["neighbor_id_2d"].copy()
    v = int(np.nonzero(nb[:, 0] >= 0)[0][len(nb) // 3])
    nb[v, 0] = (nb[v, 0] + 101) % len(nb)    # +x neighbour: a voxel elsewhere on the shell
    w = _with_table(w0, nb)
    ref = run_oracle(w, max_beats=7)
    prep = _prep(w)
    plain = runtime.PlainSim(prep, tol=1e-15)
    plain.run()
    err = np.max(np.abs(plain.su.cpu().numpy() - ref["u"])) / np.max(np.abs(ref["u"]))
    assert err <= 1e-9
    assert np.array_equal(plain.lat.cpu().numpy().astype(np.int64), ref["lat"][:7])
    arena = runtime.Arena(prep.n_voxel, 1, runtime.SolverConfig())
    with pytest.raises(_lib.PhsError, match="not the adjacent lattice voxel"):
        arena.add(prep)
