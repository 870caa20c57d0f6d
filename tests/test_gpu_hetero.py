"""Per-voxel cell parameters (P_2d[:, 15:20] varying over voxels) on the device, against the
FP64 oracle, through the C ABI.

The reference's reaction kernel reads every parameter per voxel (simulation/simulation_gpu.py:
29-41, packed by simulation/equation_parts.py:99-110).  Here the parameters vary smoothly
across each shell (Mitchell-Schaeffer tau_in, tau_out, tau_open, tau_close; Aliev-Panfilov k,
eps0, mu1).  Bars: plain FP64 within 1e-9 of the oracle with identical activation times; FP32
arena within 1 step of the oracle's activation times with the same voxels activated; a
heterogeneous simulation packed with a uniform one gives each the same bits as alone.
"""
import numpy as np
import pytest

from tests.helpers import run_oracle, small_shell

pytestmark = pytest.mark.gpu


def _hetero(model):
    from pyheartsim_amd import configs
    w = small_shell(t_final=40.0, heart_model_flag=model, seed=6)
    g = w["geometry"]
    x = g["voxel"][:, 0]
    s = (x - x.min()) / (x.max() - x.min())    # 0 .. 1 across the shell
    hm = dict(w["hm"])
    if model == 0:
        hm["tau_in_voxel"] = 0.25 + 0.1 * s
        hm["tau_out_voxel"] = 5.0 + 2.0 * s
        hm["tau_open_voxel"] = 100.0 + 40.0 * s
        hm["tau_close_voxel"] = 80.0 + 70.0 * s
    else:
        hm["k_voxel"] = 7.0 + 2.0 * s
        hm["epsilon_0_voxel"] = 0.002 + 0.002 * s
        hm["mu1_voxel"] = 0.15 + 0.1 * s
    w["hm"] = hm
    w["P_2d"] = configs.build_P_2d(g, w["sim"], hm)
    return w


@pytest.mark.parametrize("model", [0, 1])
def test_per_voxel_parameters_against_oracle(cuda, model):
    from pyheartsim_amd import runtime
    w = _hetero(model)
    P = w["P_2d"]
    assert np.ptp(P[:, 15]) > 0
    ref = run_oracle(w, max_beats=7)
    prep = runtime.prepare(P.shape[0], P, w["geometry"], w["sim"], w["arr"])
    assert prep.params_voxel is not None
    plain = runtime.PlainSim(prep, tol=1e-15)
    plain.run()
    err = np.max(np.abs(plain.su.cpu().numpy() - ref["u"])) / np.max(np.abs(ref["u"]))
    print(f"model {model} N={prep.n_voxel} plain rel err {err:.2e}")
    assert err <= 1e-9
    assert np.array_equal(plain.lat.cpu().numpy().astype(np.int64), ref["lat"][:7])
    arena = runtime.Arena(prep.n_voxel, 1, runtime.SolverConfig())
    slot = arena.add(prep)
    arena.step(prep.n_steps)
    lat = arena.lat(slot).cpu().numpy().astype(np.int64)
    r = ref["lat"][:7]
    assert np.array_equal(lat >= 0, r >= 0)
    m = r >= 0
    assert m.sum() > 0.25 * prep.n_voxel
    d = int(np.max(np.abs(lat[m] - r[m])))
    print(f"fp32 LAT max diff {d} over {int(m.sum())} activations")
    assert d <= 1
    # the heterogeneous parameters matter: the uniform run activates differently
    uni = run_oracle(small_shell(t_final=40.0, heart_model_flag=model, seed=6), max_beats=7)
    assert not np.array_equal(uni["lat"][:7], r)


def test_hetero_packed_with_uniform_equals_alone(cuda):
    from pyheartsim_amd import runtime
    wh = _hetero(0)
    wu = small_shell(t_final=30.0, seed=3)
    ph = runtime.prepare(wh["P_2d"].shape[0], wh["P_2d"], wh["geometry"], wh["sim"], wh["arr"])
    pu = runtime.prepare(wu["P_2d"].shape[0], wu["P_2d"], wu["geometry"], wu["sim"], wu["arr"])
    alone = []
    for p in (ph, pu):
        a = runtime.Arena(p.n_voxel, 1, runtime.SolverConfig())
        s = a.add(p)
        a.step(p.n_steps)
        alone.append(a.outputs(s))
        a.close()
    mix = runtime.Arena(ph.n_voxel + pu.n_voxel, 2, runtime.SolverConfig())
    su = mix.add(pu)
    mix.step(40)
    sh = mix.add(ph)   # the arena's per-voxel parameter arrays appear mid-run
    for _ in range(ph.n_steps // 50 + 1):
        mix.step(50, graph=True)
    for s, (ua, _, la) in zip((sh, su), alone):
        um, _, lm = mix.outputs(s)
        assert np.array_equal(ua.cpu().numpy(), um.cpu().numpy())
        assert np.array_equal(la.cpu().numpy(), lm.cpu().numpy())
