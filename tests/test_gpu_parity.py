"""GPU parity of the simulation hot path against the FP64 oracle (through the C ABI).

Bars (written in each test):
  * plain FP64 form: max |dVm| / max |Vm_oracle| <= 1e-9, LAT identical;
  * bricked FP32 form: LAT within 1 dt of the oracle for every voxel and beat;
  * alone vs packed with other simulations: bitwise identical.
"""
import numpy as np
import pytest

from tests.helpers import run_oracle, small_sheet, small_shell

pytestmark = pytest.mark.gpu


def _prep(w):
    from pyheartsim_amd import runtime
    P = w["P_2d"]
    return runtime.prepare(P.shape[0], P, w["geometry"], w["sim"], w["arr"])


def _plain(w, probes=None):
    from pyheartsim_amd import runtime
    prep = _prep(w)
    sim = runtime.PlainSim(prep, tol=1e-15, probes=probes)
    sim.run()
    return prep, sim


def _lat_ok(lat_a, lat_b, tol_steps):
    a = np.asarray(lat_a, dtype=np.int64)
    b = np.asarray(lat_b, dtype=np.int64)
    assert np.array_equal(a >= 0, b >= 0), "different voxels activated"
    m = a >= 0
    return int(np.max(np.abs(a[m] - b[m]))) if m.any() else 0


@pytest.mark.parametrize("scar", [0, 2])
def test_plain_fp64_matches_oracle_shell(cuda, scar):
    w = small_shell(n_scar=scar, t_final=40.0)
    ref = run_oracle(w, max_beats=7)
    prep, sim = _plain(w)
    print(f"shell N={prep.n_voxel} steps={prep.n_steps} sweeps={sim.info()}", flush=True)
    su = sim.su.cpu().numpy()
    T = ref["u"].shape[0]
    assert su.shape[0] == T
    err = np.max(np.abs(su - ref["u"])) / np.max(np.abs(ref["u"]))
    print("rel Vm err", err)
    assert err <= 1e-9
    sh = sim.sh.cpu().numpy()
    assert np.max(np.abs(sh - ref["h"])) <= 1e-9
    assert np.array_equal(sim.lat.cpu().numpy().astype(np.int64), ref["lat"][:7])


def test_plain_fp64_matches_oracle_sheet_with_temp_block(cuda):
    w = small_sheet(n=32, n_steps=400)
    nf = np.zeros(w["P_2d"].shape[0], dtype=np.int64)
    ijk = w["geometry"]["ijk"]
    nf[(ijk[:, 0] == 20) & (ijk[:, 1] < 20)] = 3
    w["arr"]["node_flag"] = nf
    probes = w["probes"]
    ref = run_oracle(w, probes=probes)
    prep, sim = _plain(w, probes=probes)
    err = np.max(np.abs(sim.su.cpu().numpy() - ref["u"])) / np.max(np.abs(ref["u"]))
    assert err <= 1e-9
    perr = np.max(np.abs(sim.probe_u.cpu().numpy() - ref["probe_u"]))
    assert perr <= 1e-9 * np.max(np.abs(ref["probe_u"]))
    assert np.array_equal(sim.lat.cpu().numpy().astype(np.int64), ref["lat"][:7])


def test_fp32_arena_lat_within_one_step(cuda):
    from pyheartsim_amd import runtime
    w = small_shell(t_final=40.0)
    ref = run_oracle(w, max_beats=7)
    prep = _prep(w)
    arena = runtime.Arena(prep.n_voxel, 1, runtime.SolverConfig())
    slot = arena.add(prep)
    arena.step(prep.n_steps)
    su, _, lat = arena.outputs(slot)
    d = _lat_ok(lat.cpu().numpy(), ref["lat"][:7], 1)
    vm = np.max(np.abs(su.cpu().numpy() - ref["u"]))
    print(f"fp32 arena: max LAT diff {d} steps, max |dVm| {vm:.3e}")
    assert d <= 1


def test_arena_alone_equals_mixed_and_graph(cuda):
    """A simulation gives the same bits alone (lean kernels: no scar, one sweep count) as
    packed with a scarred one (lean kernels with the scar test), admitted mid-run, with
    graph replay.  The identity-sweep flavours' bits (lean, lean with the block test,
    general) are checked the same way by test_per_simulation_sweep_counts_mixed_spacing."""
    from pyheartsim_amd import runtime
    wa = small_shell(t_final=20.0, seed=3)
    wb = small_shell(t_final=30.0, seed=5, semi_axes=(8.0, 7.0, 6.0), n_scar=1)
    pa, pb = _prep(wa), _prep(wb)
    solver = runtime.SolverConfig(sweeps=12)
    alone = runtime.Arena(pa.n_voxel, 1, solver)
    s = alone.add(pa)
    alone.step(pa.n_steps)
    ua, _, la = alone.outputs(s)
    mixed = runtime.Arena(pa.n_voxel + pb.n_voxel, 2, solver)
    sb = mixed.add(pb)
    mixed.step(100)
    sa = mixed.add(pa)  # admitted while the other simulation is mid-run
    for _ in range(pa.n_steps // 50 + 1):
        mixed.step(50, graph=True)
    um, _, lm = mixed.outputs(sa)
    assert np.array_equal(ua.cpu().numpy(), um.cpu().numpy())
    assert np.array_equal(la.cpu().numpy(), lm.cpu().numpy())


@pytest.mark.parametrize("removed", [0, 1])
def test_compaction_moves_records_bitwise(cuda, removed):
    """Removing a simulation and compacting slides the ones above it down in brick order:
    with the bottom one removed, the small middle one moves by more than its size and the
    large top one onto part of its old range; with the small middle one removed, the large
    top one moves onto part of its old range.  The launches shrink, and the simulations that
    stay give the same bits as alone, with graph replay across the move."""
    from pyheartsim_amd import runtime
    axes = [(9.0, 8.0, 7.0), (6.0, 5.0, 5.0), (10.0, 9.0, 8.0)]
    ws = [small_shell(t_final=10.0 if i == removed else 30.0, seed=3 + 2 * i, semi_axes=ax,
                      n_scar=1 if i == 2 else 0) for i, ax in enumerate(axes)]
    ps = [_prep(w) for w in ws]
    solver = runtime.SolverConfig()
    ref = {}
    for i, p in enumerate(ps):
        if i == removed:
            continue
        one = runtime.Arena(p.n_voxel, 1, solver)
        s = one.add(p)
        one.step(p.n_steps)
        ref[i] = [t.cpu().numpy() for t in one.outputs(s) if t is not None]
        one.close()
    ar = runtime.Arena(sum(p.n_voxel for p in ps), 3, solver)
    slots = [ar.add(p) for p in ps]
    nb = [ar.slot_info(s)["n_bricks"] for s in slots]
    assert nb[1] < nb[0] < nb[2]
    ar.step(ps[removed].n_steps, graph=True)
    before = ar.counts()
    ar.remove(slots[removed])
    assert ar.compact(0.0) >= 1
    after = ar.counts()
    assert after["bricks"] == before["bricks"] - nb[removed]
    assert after["launch_bricks"] < before["launch_bricks"]
    assert after["launch_bricks"] - after["bricks"] < 8 * 4   # contiguous up to the grid's rounding
    assert ar.compact(0.0) == 0                                # nothing left to move
    for _ in range(max(p.n_steps for p in ps) // 50 + 1):
        ar.step(50, graph=True)
    for i, s in enumerate(slots):
        if i == removed:
            continue
        got = [t.cpu().numpy() for t in ar.outputs(s) if t is not None]
        assert len(got) == len(ref[i]) == 2      # samples and activation times
        for g, r in zip(got, ref[i]):
            assert np.array_equal(g, r)
    ar.close()


@pytest.mark.parametrize("refine", [0, 1])
def test_voxel_compaction_admits_into_freed_ranges(cuda, refine):
    """An arena exactly as large as three simulations: with the outer two removed, a fourth
    that fits in neither freed voxel range is refused (PHS_ERR_CAPACITY).  Compacting with
    voxels moves the middle one's voxel data down, the fourth then fits, and both compute
    bitwise what they compute alone (refine 1: with the FP64 refinement accumulator and the
    diagnostic u* arrays, which move too; the moved one's CN residual is recomputed)."""
    from pyheartsim_amd import _lib, runtime
    axes = {"a": (6.5, 5.5, 5.0), "b": (9.0, 8.0, 7.0), "c": (7.0, 6.0, 5.0),
            "d": (7.5, 6.5, 5.5)}
    ps = {k: _prep(small_shell(t_final=10.0 if k in "ac" else 30.0, seed=3 + i, semi_axes=ax))
          for i, (k, ax) in enumerate(axes.items())}
    n = {k: p.n_voxel for k, p in ps.items()}
    assert max(n["a"], n["c"]) < n["d"] <= n["a"] + n["c"]
    solver = runtime.SolverConfig(refine=1, refine_sweeps=8, diag=True) if refine else \
        runtime.SolverConfig()
    ref = {}
    for k in "bd":
        one = runtime.Arena(n[k], 1, solver)
        s = one.add(ps[k])
        one.step(ps[k].n_steps)
        ref[k] = [t.cpu().numpy() for t in one.outputs(s) if t is not None]
        one.close()
    ar = runtime.Arena(n["a"] + n["b"] + n["c"], 3, solver)
    sa, sb, sc = ar.add(ps["a"]), ar.add(ps["b"]), ar.add(ps["c"])
    ar.step(ps["a"].n_steps, graph=True)
    ar.remove(sa)
    ar.remove(sc)
    assert ar.free_capacity()[0] >= n["d"]
    with pytest.raises(_lib.PhsError) as e:
        ar.add(ps["d"])
    assert e.value.code == _lib.PHS_ERR_CAPACITY
    assert ar.compact(0.0, voxels=True) >= 1
    assert ar.slot_info(sb)["voxel_offset"] == 0
    sd = ar.add(ps["d"])
    for _ in range(ps["d"].n_steps // 50 + 1):
        ar.step(50, graph=True)
    for k, s in (("b", sb), ("d", sd)):
        got = [t.cpu().numpy() for t in ar.outputs(s) if t is not None]
        assert len(got) == len(ref[k]) == 2
        for g, r in zip(got, ref[k]):
            assert np.array_equal(g, r), k
    if refine:
        assert ar.last_residual(sb) <= 1e-6   # recomputed from the moved u* and correction
    ar.close()


def test_fp32_arena_temp_block_and_scar_vs_oracle(cuda):
    from pyheartsim_amd import runtime
    w = small_shell(t_final=40.0, n_scar=2, temp_block=True)
    ref = run_oracle(w, max_beats=7)
    prep = _prep(w)
    assert (prep.vflags & 1).any() and (prep.vflags & 2).any()
    arena = runtime.Arena(prep.n_voxel, 1, runtime.SolverConfig())
    slot = arena.add(prep)
    arena.step(prep.n_steps)
    _, _, lat = arena.outputs(slot)
    assert _lat_ok(lat.cpu().numpy(), ref["lat"][:7], 1) <= 1


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_dropin_compute_matches_oracle(cuda, precision):
    """simulation_gpu.compute keeps the reference signature and return layout."""
    from pyheartsim_amd.simulation import simulation_gpu
    w = small_shell(t_final=30.0, n_scar=1)
    P = w["P_2d"]
    ref = run_oracle(w, max_beats=7)
    u, h, t, lat = simulation_gpu.compute(P.shape[0], P, w["geometry"], w["sim"], w["arr"],
                                          precision=precision, return_lat=True)
    assert u.dtype == np.float32 and h.dtype == np.float32 and t.dtype == np.float32
    assert u.shape == ref["u"].shape and h.shape == ref["h"].shape
    assert np.array_equal(t, ref["physical_time"])
    if precision == "fp64":
        # outputs are float32 (reference dtype): the bar is FP32 rounding of the FP64 state
        assert np.max(np.abs(u - ref["u"])) <= 1e-6
        assert np.array_equal(lat.astype(np.int64), ref["lat"][:7])
    else:
        assert _lat_ok(lat, ref["lat"][:7], 1) <= 1
        assert np.max(np.abs(u - ref["u"])) <= 5e-2


@pytest.mark.parametrize("flavour", ["lean", "flags", "general"])
def test_per_simulation_sweep_counts_mixed_spacing(cuda, flavour):
    """Automatic sweeps: a 0.25 mm shell (lambda_max ~ 20) needs more sweeps than a 0.5 mm
    one (~ 6).  Packed together the arena launches the larger count, and each simulation
    still computes bitwise what it computes alone (identity sweeps past its own count);
    the fine one stays within 1 step of the FP64 oracle's LAT.  The coarse shell has no
    temporary block (lean kernels with identity sweeps), a temporary block (the same with
    the block test) or per-voxel conductivity (general kernels)."""
    from pyheartsim_amd import runtime
    coarse = small_shell(t_final=20.0, seed=3, temp_block=flavour == "flags")
    if flavour == "general":
        coarse = _hetero_c(coarse)
    fine = small_shell(delta=0.25, semi_axes=(6.0, 5.5, 5.0), wall=1.5, t_final=20.0, seed=4)
    pc, pf = _prep(coarse), _prep(fine)
    solver = runtime.SolverConfig()           # sweeps = 0: automatic
    res = {}
    for name, p in (("coarse", pc), ("fine", pf)):
        a = runtime.Arena(p.n_voxel, 1, solver)
        s = a.add(p)
        lmax = a.bounds(s)[1]
        assert a.sweeps() == runtime.stable_sweeps(lmax, solver.guess)[1]
        a.step(p.n_steps)
        res[name] = (a.sweeps(), [t.cpu().numpy() for t in a.outputs(s)[::2]])
        a.close()
    assert res["fine"][0] > res["coarse"][0]
    mix = runtime.Arena(pc.n_voxel + pf.n_voxel, 2, solver)
    sf = mix.add(pf)
    sc = mix.add(pc)
    assert mix.sweeps() == res["fine"][0]
    for _ in range(max(pc.n_steps, pf.n_steps) // 50 + 1):
        mix.step(50, graph=True)
    for name, s in (("coarse", sc), ("fine", sf)):
        um, _, lm = mix.outputs(s)
        ua, la = res[name][1]
        assert np.array_equal(ua, um.cpu().numpy()), name
        assert np.array_equal(la, lm.cpu().numpy()), name
    ref = run_oracle(fine, max_beats=7)
    assert _lat_ok(res["fine"][1][1], ref["lat"][:7], 1) <= 1


def test_fixed_sweeps_below_stability_limit_rejected(cuda):
    from pyheartsim_amd import _lib, runtime
    fine = small_shell(delta=0.25, semi_axes=(6.0, 5.5, 5.0), wall=1.5, t_final=5.0, seed=4)
    p = _prep(fine)
    a = runtime.Arena(p.n_voxel, 1, runtime.SolverConfig(sweeps=4))
    with pytest.raises(_lib.PhsError, match="unstable"):
        a.add(p)


def _run_pair_with_removal(solver, wa, wb):
    """Two simulations packed; the first is removed after 60 steps (its bricks stay in the
    table without a slot), the second runs to its end with graph replays."""
    from pyheartsim_amd import runtime
    pa, pb = _prep(wa), _prep(wb)
    a = runtime.Arena(pa.n_voxel + pb.n_voxel, 2, solver)
    sa = a.add(pa)
    sb = a.add(pb)
    a.step(60)
    a.remove(sa)
    while a.steps_done(sb) < pb.n_steps:
        a.step(50, graph=True)
    u, _, lat = a.outputs(sb)
    out = (u.cpu().numpy(), lat.cpu().numpy())
    a.close()
    return out


@pytest.mark.parametrize("brick,guess,refine", [((4, 4, 4), 1, 0), ((8, 8, 1), 2, 0),
                                                ((4, 4, 4), 0, 1)])
def test_two_bricks_per_wavefront_equals_one(cuda, brick, guess, refine):
    """Two bricks per wavefront (bricks_per_wave = 2) compute every voxel with the same
    operations in the same order as one brick per wavefront: bitwise equal, also with
    removed simulations' bricks (no slot) in the table."""
    from pyheartsim_amd import runtime
    wa = small_shell(t_final=10.0, seed=3)
    wb = small_shell(t_final=20.0, seed=5, semi_axes=(8.0, 7.0, 6.0), n_scar=1, temp_block=True)
    res = {}
    for bpw in (1, 2):
        solver = runtime.SolverConfig(brick=brick, guess=guess, refine=refine, bricks_per_wave=bpw)
        res[bpw] = _run_pair_with_removal(solver, wa, wb)
    assert np.array_equal(res[2][0], res[1][0])
    assert np.array_equal(res[2][1], res[1][1])


@pytest.mark.parametrize("bpw", [1, 2])
def test_graph_recaptured_when_bricks_added_within_grid(cuda, bpw):
    """A simulation admitted between graph replays is stepped by the next replay also when
    the launch grid does not change (a few bricks added below the grid's capacity)."""
    from pyheartsim_amd import runtime
    pa, pb = _prep(small_sheet(n=24, n_steps=200)), _prep(small_sheet(n=8, n_steps=200))
    solver = runtime.SolverConfig(bricks_per_wave=bpw)
    alone = runtime.Arena(pb.n_voxel, 1, solver)
    s = alone.add(pb)
    alone.step(pb.n_steps)
    ua, _, la = alone.outputs(s)
    mixed = runtime.Arena(pa.n_voxel + pb.n_voxel, 2, solver)
    mixed.add(pa)
    mixed.step(50, graph=True)
    g0 = mixed.counts()["launch_bricks"]
    sb = mixed.add(pb)
    print(f"bricks {mixed.counts()['bricks']} launch bricks {g0} -> {mixed.counts()['launch_bricks']}")
    while mixed.steps_done(sb) < pb.n_steps:
        mixed.step(50, graph=True)
    um, _, lm = mixed.outputs(sb)
    assert np.array_equal(ua.cpu().numpy(), um.cpu().numpy())
    assert np.array_equal(la.cpu().numpy(), lm.cpu().numpy())


@pytest.mark.parametrize("scar", [False, True])
def test_odd_graph_lengths_and_steps_past_the_end_equal_eager(cuda, scar):
    """The per-step slot values alternate between two copies with the launch parity.  Graph
    replays of odd lengths (both starting parities), eager steps in between, a simulation
    admitted at an odd parity, and steps run after both simulations have finished give the
    same bits as plain eager stepping, and a finished simulation stays finished.  With a
    scarred second simulation the arena runs the general kernels, otherwise the lean ones."""
    from pyheartsim_amd import runtime
    pa = _prep(small_sheet(n=16, n_steps=120))
    pb = (_prep(small_shell(t_final=6.0, seed=5, semi_axes=(7.0, 6.0, 5.0), n_scar=1)) if scar
          else _prep(small_sheet(n=12, n_steps=90)))
    ref = {}
    for name, p in (("a", pa), ("b", pb)):
        alone = runtime.Arena(p.n_voxel, 1, runtime.SolverConfig())
        s = alone.add(p)
        alone.step(p.n_steps)
        u, _, lat = alone.outputs(s)
        ref[name] = (u.cpu().numpy(), lat.cpu().numpy())
    mixed = runtime.Arena(pa.n_voxel + pb.n_voxel, 2, runtime.SolverConfig())
    sa = mixed.add(pa)
    mixed.step(7, graph=True)
    mixed.step(3)
    sb = mixed.add(pb)  # admitted after 10 steps, then 7-step replays from the other parity
    for n, graph in ((7, True), (13, True), (7, True), (1, False), (7, True), (13, True)):
        mixed.step(n, graph=graph)
    while mixed.steps_done(sa) < pa.n_steps or mixed.steps_done(sb) < pb.n_steps:
        mixed.step(7, graph=True)
    mixed.step(13, graph=True)  # past the end of both
    mixed.step(2)
    for name, s in (("a", sa), ("b", sb)):
        u, _, lat = mixed.outputs(s)
        print(f"sim {name}: samples {u.shape}", flush=True)
        assert np.array_equal(ref[name][0], u.cpu().numpy())
        assert np.array_equal(ref[name][1], lat.cpu().numpy())


def _hetero_c(w, seed=11):
    """Per-voxel conductivity (P_2d column 20, as the reference builds it from c_voxel) varying
    smoothly in [0.6, 1.4] x the configured value, with one region at 0 (scar by conductivity)."""
    rng = np.random.default_rng(seed)
    P = w["P_2d"].copy()
    v = w["geometry"]["voxel"]
    k = rng.normal(size=3)
    f = 0.6 + 0.8 * (0.5 + 0.5 * np.sin(v @ k / 3.0))
    P[:, 20] = P[:, 20] * f
    P[np.linalg.norm(v - v[len(v) // 3], axis=1) < 1.2, 20] = 0.0
    w["P_2d"] = P
    return w


def test_plain_fp64_matches_oracle_heterogeneous_c(cuda):
    w = _hetero_c(small_shell(t_final=40.0))
    ref = run_oracle(w, max_beats=7)
    prep, sim = _plain(w)
    assert prep.c_voxel is not None
    err = np.max(np.abs(sim.su.cpu().numpy() - ref["u"])) / np.max(np.abs(ref["u"]))
    print("rel Vm err", err)
    assert err <= 1e-9
    assert np.array_equal(sim.lat.cpu().numpy().astype(np.int64), ref["lat"][:7])


def test_fp32_arena_heterogeneous_c_lat_within_one_step(cuda):
    from pyheartsim_amd import runtime
    w = _hetero_c(small_shell(t_final=40.0))
    ref = run_oracle(w, max_beats=7)
    prep = _prep(w)
    arena = runtime.Arena(prep.n_voxel, 1, runtime.SolverConfig())
    slot = arena.add(prep)
    arena.step(prep.n_steps)
    _, _, lat = arena.outputs(slot)
    assert _lat_ok(lat.cpu().numpy(), ref["lat"][:7], 1) <= 1


def test_aliev_panfilov_shell_against_oracle(cuda):
    """The 3D Aliev-Panfilov path (time scaling by 6, AP reaction): plain FP64 form within
    1e-9 relative Vm of the oracle with identical LAT, the bricked FP32 form within 1 step."""
    from pyheartsim_amd import runtime
    w = small_shell(t_final=80.0, heart_model_flag=1)
    ref = run_oracle(w, max_beats=7)
    assert (ref["lat"][0] >= 0).mean() > 0.9     # the wave crosses the shell
    prep, sim = _plain(w)
    assert prep.model == 1
    err = np.max(np.abs(sim.su.cpu().numpy() - ref["u"])) / np.max(np.abs(ref["u"]))
    print("AP rel Vm err", err)
    assert err <= 1e-9
    assert np.array_equal(sim.lat.cpu().numpy().astype(np.int64), ref["lat"][:7])
    arena = runtime.Arena(prep.n_voxel, 1, runtime.SolverConfig())
    slot = arena.add(prep)
    arena.step(prep.n_steps)
    _, _, lat = arena.outputs(slot)
    assert _lat_ok(lat.cpu().numpy(), ref["lat"][:7], 1) <= 1
