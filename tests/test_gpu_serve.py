"""Continuous batching: requests admitted while others run give the same results as alone.

Bar: the electrogram and activation times of a request served in a mix of concurrent,
differently sized requests are bitwise identical to the same request served alone, and
its LAT is within 1 step of the plain FP64 form.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_serve_mix_equals_alone_and_lat(cuda):
    from pyheartsim_amd import runtime, serve
    reqs = serve.make_requests(6, qps=200.0, seed=5, n_range=(4e3, 2.5e4),
                               duration_range=(40.0, 90.0), electrode_range=(8, 40))
    srv = serve.Server(max_voxels=400_000, max_sims=8, hbm_budget_bytes=8e9, ring=16,
                       chunk_steps=50, keep_results=True)
    st = srv.run(reqs)
    assert st.completed == len(reqs)
    assert all(l > 0 for l in st.latencies) and st.p99 >= st.p50
    by_id = {r["rid"]: r for r in st.records}
    probe = reqs[3]
    alone = serve.Server(max_voxels=100_000, max_sims=1, hbm_budget_bytes=8e9, ring=16,
                         chunk_steps=50, keep_results=True)
    probe_alone = serve.Request(**{**probe.__dict__, "arrival": 0.0})
    sa = alone.run([probe_alone])
    ra = sa.records[0]
    rm = by_id[probe.rid]
    assert np.array_equal(ra["egm"].numpy(), rm["egm"].numpy())
    assert np.array_equal(ra["lat"].numpy(), rm["lat"].numpy())
    # LAT of the served request against the plain FP64 form of the same simulation
    plain = runtime.PlainSim(probe.prep, tol=1e-15, samples=False)
    plain.run()
    ref = plain.lat.cpu().numpy()
    got = rm["lat"].numpy()
    assert np.array_equal(got >= 0, ref >= 0)
    m = ref >= 0
    assert np.max(np.abs(got[m].astype(np.int64) - ref[m])) <= 1
    egm = rm["egm"].numpy()
    assert egm.shape == (probe.prep.n_saved(), len(probe.electrodes))
    assert np.isclose(np.mean(np.ptp(egm, axis=0)), 1.0, rtol=1e-5)


def test_ragged_batch_matches_alone(cuda):
    """Uneven simulations with scar packed in one launch sequence = each alone (bitwise)."""
    from pyheartsim_amd import runtime
    from tests.helpers import small_shell
    ws = [small_shell(t_final=25.0, seed=s, semi_axes=ax, n_scar=ns)
          for s, ax, ns in ((3, (9.0, 8.0, 7.0), 0), (4, (7.0, 6.0, 5.0), 2), (8, (11.0, 9.0, 8.0), 1))]
    preps = [runtime.prepare(w["P_2d"].shape[0], w["P_2d"], w["geometry"], w["sim"], w["arr"])
             for w in ws]
    solver = runtime.SolverConfig()
    alone = []
    for p in preps:
        a = runtime.Arena(p.n_voxel, 1, solver)
        s = a.add(p)
        a.step(p.n_steps)
        alone.append(a.outputs(s))
    mix = runtime.Arena(sum(p.n_voxel for p in preps), len(preps), solver)
    slots = [mix.add(p) for p in preps]
    for _ in range(max(p.n_steps for p in preps) // 50 + 1):
        mix.step(50, graph=True)
    for s, (ua, _, la) in zip(slots, alone):
        um, _, lm = mix.outputs(s)
        assert np.array_equal(ua.cpu().numpy(), um.cpu().numpy())
        assert np.array_equal(la.cpu().numpy(), lm.cpu().numpy())


def test_full_arena_reports_capacity_and_server_queues(cuda):
    """An add that does not fit fails with PHS_ERR_CAPACITY (slots, then voxels), and the
    server keeps such a request queued until a running one finishes."""
    from pyheartsim_amd import _lib, runtime, serve
    from tests.helpers import small_shell
    w = small_shell(t_final=10.0)
    p = runtime.prepare(w["P_2d"].shape[0], w["P_2d"], w["geometry"], w["sim"], w["arr"])
    a = runtime.Arena(4 * p.n_voxel, 1, runtime.SolverConfig())
    a.add(p)
    with pytest.raises(_lib.PhsError) as e:
        a.add(p)
    assert e.value.code == _lib.PHS_ERR_CAPACITY
    a.close()
    b = runtime.Arena(p.n_voxel // 2, 2, runtime.SolverConfig())
    with pytest.raises(_lib.PhsError) as e:
        b.add(p)
    assert e.value.code == _lib.PHS_ERR_CAPACITY
    b.close()
    reqs = serve.make_requests(3, qps=1000.0, seed=9, n_range=(4e3, 8e3),
                               duration_range=(20.0, 30.0), electrode_range=(8, 16))
    srv = serve.Server(max_voxels=20_000, max_sims=1, hbm_budget_bytes=8e9, ring=16,
                       chunk_steps=50)
    st = srv.run(reqs)
    assert st.completed == len(reqs)


def test_server_compacts_voxels_to_admit(cuda):
    """A served arena exactly as large as its first three requests: once the short outer two
    finish, the fourth fits in neither freed voxel range, so the server compacts the voxel
    ranges and admits it while the long middle one still runs.  The fourth's electrogram and
    activation times equal those it gets served alone, bit for bit."""
    from pyheartsim_amd import serve

    def req(rid, n_target, seed, duration):
        return serve.build_request(dict(rid=rid, arrival=0.0, n_target=n_target, seed=seed,
                                        model=0, duration_ms=duration, n_electrodes=8,
                                        delta=0.4, dt_ms=0.05))
    a, b, c, d = req(0, 2e4, 2, 10.0), req(1, 6e4, 4, 60.0), req(2, 2.5e4, 6, 10.0), \
        req(3, 3e4, 3, 20.0)
    n = [r.prep.n_voxel for r in (a, b, c, d)]
    assert max(n[0], n[2]) < n[3] <= n[0] + n[2]
    srv = serve.Server(max_voxels=n[0] + n[1] + n[2], max_sims=4, hbm_budget_bytes=8e9,
                       ring=16, chunk_steps=50, keep_results=True)
    st = srv.run([a, b, c, d])
    assert st.completed == 4
    rec = {r["rid"]: r for r in st.records}
    assert rec[3]["admitted"] < rec[1]["done"]     # admitted beside the running middle one
    assert st.voxel_moves >= 1
    alone = serve.Server(max_voxels=n[3], max_sims=1, hbm_budget_bytes=8e9, ring=16,
                         chunk_steps=50, keep_results=True)
    ra = alone.run([serve.Request(**{**d.__dict__, "rid": 9})]).records[0]
    assert np.array_equal(ra["egm"].numpy(), rec[3]["egm"].numpy())
    assert np.array_equal(ra["lat"].numpy(), rec[3]["lat"].numpy())
