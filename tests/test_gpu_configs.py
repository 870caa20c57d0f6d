"""The measured workloads at their real sizes on the GPU (through the C ABI).

  shell      the 0.4 mm Mitchell-Schaeffer shell (199 446 voxels), 2000 ms, 64 electrodes
  precision  the 0.2 mm shell with 3 scar blobs (15 % of the wall, 1.6 M voxels), FP64-residual
             refinement, 256 electrodes
  batched    members of the 64-shell batched workload (0.4 mm, 0-3 scar blobs, CL 250-600 ms),
             packed and alone, with streamed electrograms
  served     requests drawn from the served workload's ranges (2e4-2e6 voxels, MS/AP,
             32-256 electrodes), durations trimmed to 40-80 ms

Oracles: the C/OpenMP FP64 restatement (oracle/cpu_oracle.c, tight BiCGSTAB; checked against
the NumPy oracle in test_cpu_oracle.py) for the shell's first 400 steps; the FP64 electrogram
oracle (oracle/egm_oracle.py) for the electrograms; the plain FP64 GPU form (itself checked
against the oracle to 1e-9 in test_gpu_parity.py and above) for full-length activation times,
where the NumPy oracle would take hours.

Bars: plain FP64 within 1e-9 of the oracle; chosen FP32 activation times within 1 step of the
plain FP64 form with the same voxels activated; electrograms within rel-L2 1e-3 of the FP64
oracle; CN residual <= 1e-10 with refinement; packed (or served in a mix) = alone, bitwise.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _identity_D(n):
    D = np.zeros((9, n))
    D[0] = D[4] = D[8] = 1.0
    return D


def _lat_within_one(lat, ref, n_steps):
    lat = np.asarray(lat, dtype=np.int64)
    ref = np.asarray(ref, dtype=np.int64)
    only = (lat >= 0) != (ref >= 0)
    # an activation in the window's last step that the other form makes one step later
    at_end = only & (np.maximum(lat, ref) == n_steps - 1)
    assert not np.any(only & ~at_end), "different voxels activated"
    m = (lat >= 0) & (ref >= 0)
    d = int(np.max(np.abs(lat[m] - ref[m]))) if m.any() else 0
    assert d <= 1, f"LAT differs by {d} steps"
    return int(m.sum())


def _step_to_end(arena, slot, n_steps, chunk=500):
    while arena.steps_done(slot) < n_steps:
        arena.step(chunk, graph=True)


@pytest.mark.timeout(600)
def test_shell_full_size(cuda):
    torch = cuda
    from oracle import cpu_oracle, egm_oracle
    from pyheartsim_amd import configs, runtime
    from pyheartsim_amd.simulation.unipolar_electrogram_gpu import LeadField
    w = configs.shell_ms()
    g = w["geometry"]
    P = configs.build_P_2d(g, w["sim"], w["hm"])
    prep = runtime.prepare(P.shape[0], P, g, w["sim"], w["arr"])
    N = prep.n_voxel
    print(f"shell N={N} steps={prep.n_steps}", flush=True)
    assert N == 199446 and prep.n_steps == 40000
    plain = runtime.PlainSim(prep, tol=1e-15, samples=True, sample_h=False, sample_fp64=False)
    plain.run(400)
    u_p = plain.state()[0].cpu().numpy()
    cs = cpu_oracle.CpuSim(P, g, w["sim"], w["arr"], rtol=1e-12, atol=0.0)
    assert cs.run(400) == 0
    assert np.max(cs.u) > 0.5, "no activation in the window"
    err = np.max(np.abs(u_p - cs.u)) / np.max(np.abs(cs.u))
    print(f"plain vs oracle after 400 steps: {err:.2e}", flush=True)
    assert err <= 1e-9
    plain.run()
    arena = runtime.Arena(N, 1, runtime.SolverConfig())
    slot = arena.add(prep, samples=True)
    _step_to_end(arena, slot, prep.n_steps)
    ref_lat = plain.lat.cpu().numpy()
    n = _lat_within_one(arena.lat(slot).cpu().numpy(), ref_lat, prep.n_steps)
    beats = int(np.sum(np.any(ref_lat >= 0, axis=1)))
    print(f"chosen LAT within 1 step over {n} activations, {beats} beats", flush=True)
    assert beats >= 4
    el, c = w["electrodes"], w["hm"]["c_voxel"]
    e_ref = egm_oracle.electrogram_raw(el, g["voxel"], _identity_D(N), c,
                                       plain.su.cpu().numpy(), g["Delta"], g["neighbor_id_2d"])
    lf = LeadField(el, g["voxel"], None, c, g["Delta"], g["neighbor_id_2d"])
    e_dev = lf.run(plain.su, precision="bf16x3", normalize=False).cpu().numpy()
    lf_c = LeadField(el, g["voxel"], None, c, g["Delta"], g["neighbor_id_2d"],
                     perm=arena.slots[slot].perm)
    e_chain = lf_c.run(arena.slots[slot].su, precision="bf16x3", normalize=False).cpu().numpy()
    r1, r2 = _rel(e_dev, e_ref), _rel(e_chain, e_ref)
    print(f"egm 64 x 2000: BF16 on plain samples {r1:.2e}, chosen chain {r2:.2e}")
    assert r1 <= 1e-3 and r2 <= 1e-3


@pytest.mark.timeout(600)
def test_precision_full_size(cuda):
    torch = cuda
    from oracle import egm_oracle
    from pyheartsim_amd import configs, runtime
    from pyheartsim_amd.simulation.unipolar_electrogram_gpu import LeadField
    w = configs.shell_fine(t_final=60.0)
    g = w["geometry"]
    P = configs.build_P_2d(g, w["sim"], w["hm"])
    prep = runtime.prepare(P.shape[0], P, g, w["sim"], w["arr"])
    N = prep.n_voxel
    scar = int(np.count_nonzero(prep.vflags & 1))
    print(f"precision N={N} scar={scar} steps={prep.n_steps}", flush=True)
    assert N > 1_500_000 and 0.10 * N < scar < 0.20 * N
    plain = runtime.PlainSim(prep, tol=1e-15, samples=True, sample_h=False, sample_fp64=False)
    plain.run()
    cfg = runtime.SolverConfig(refine=3, refine_sweeps=12, diag=True)
    arena = runtime.Arena(N, 1, cfg)
    slot = arena.add(prep, samples=True)
    lmax = arena.bounds(slot)[1]
    _step_to_end(arena, slot, prep.n_steps)
    res = arena.last_residual(slot)
    n = _lat_within_one(arena.lat(slot).cpu().numpy(), plain.lat.cpu().numpy(), prep.n_steps)
    print(f"lambda_max {lmax:.1f}, CN residual {res:.2e}, LAT within 1 over {n}", flush=True)
    assert res <= 1e-10 and n > 0.3 * N
    # 256 electrodes on a 32-sample block after the activation starts, against the oracle
    rows = slice(20, 52)
    block = plain.su[rows].contiguous()
    el, c = w["electrodes"], w["hm"]["c_voxel"]
    assert el.shape[0] == 256
    ap = block.cpu().numpy()
    D = _identity_D(N)
    e_ref = np.concatenate([egm_oracle.electrogram_raw(el[k:k + 32], g["voxel"], D, c, ap,
                                                       g["Delta"], g["neighbor_id_2d"])
                            for k in range(0, 256, 32)], axis=1)
    lf = LeadField(el, g["voxel"], None, c, g["Delta"], g["neighbor_id_2d"])
    e_dev = lf.run(block, precision="bf16x3", normalize=False).cpu().numpy()
    lf_c = LeadField(el, g["voxel"], None, c, g["Delta"], g["neighbor_id_2d"],
                     perm=arena.slots[slot].perm)
    e_chain = lf_c.run(arena.slots[slot].su[rows].contiguous(), precision="bf16x3",
                       normalize=False).cpu().numpy()
    r1, r2 = _rel(e_dev, e_ref), _rel(e_chain, e_ref)
    print(f"egm 256 x 32: BF16 on plain samples {r1:.2e}, chosen chain {r2:.2e}")
    assert r1 <= 1e-3 and r2 <= 1e-3


def _batched_members(indices, t_final):
    from pyheartsim_amd import configs, runtime
    kws = configs.ragged_params(64)
    out = []
    for i in indices:
        kw = dict(kws[i], t_final=t_final)
        w = configs.shell_ms(**kw)
        g = w["geometry"]
        P = configs.build_P_2d(g, w["sim"], w["hm"])
        out.append(dict(prep=runtime.prepare(P.shape[0], P, g, w["sim"], w["arr"]),
                        voxel=g["voxel"], c=w["hm"]["c_voxel"], delta=g["Delta"],
                        electrodes=w["electrodes"], n_scar=kw["n_scar"]))
    return out


def _run_chain(torch, members, chunk=100, ring=16):
    """Members packed in one arena; the rows each chunk produces are contracted into the
    electrograms through the streamed contraction.  Returns [(lat, egm)] in caller order."""
    from pyheartsim_amd import runtime
    from pyheartsim_amd.simulation.unipolar_electrogram_gpu import LeadField, StreamContraction
    arena = runtime.Arena(sum(m["prep"].n_voxel for m in members), len(members),
                          runtime.SolverConfig())
    slots = [arena.add(m["prep"], samples=True, ring=ring) for m in members]
    lfs = [LeadField(m["electrodes"], m["voxel"], None, m["c"], m["delta"], m["prep"].nb,
                     perm=arena.slots[s].perm) for m, s in zip(members, slots)]
    egms = [torch.empty((m["prep"].n_saved(), len(m["electrodes"])), device="cuda")
            for m in members]
    done = [0] * len(members)
    sc = StreamContraction()
    steps = 0
    n_steps = max(m["prep"].n_steps for m in members)
    while steps < n_steps:
        arena.step(chunk, graph=True)
        steps += chunk
        jobs = []
        for i, m in enumerate(members):
            p = m["prep"]
            upto = min(p.n_saved(), (min(steps, p.n_steps) + p.steps_per_sample - 1)
                       // p.steps_per_sample)
            su = arena.slots[slots[i]].su
            while done[i] < upto:
                r0 = done[i] % ring
                n = min(upto - done[i], ring - r0, 16)
                jobs.append((lfs[i], su[r0:r0 + n], egms[i][done[i]:done[i] + n]))
                done[i] += n
        sc.run(jobs)
    out = [(arena.lat(s).cpu().numpy(), e.cpu().numpy()) for s, e in zip(slots, egms)]
    for lf in lfs:
        lf.close()
    arena.close()
    return out


@pytest.mark.timeout(600)
def test_batched_members_packed_equal_alone(cuda):
    torch = cuda
    from pyheartsim_amd import runtime
    # 24 497 bricks packed (the record prefetch is on from 16 384 bricks), 7 938 / 5 266 /
    # 11 293 alone (off): the bitwise comparison also covers the prefetch
    members = _batched_members([0, 2, 49], t_final=200.0)
    ns = [m["prep"].n_voxel for m in members]
    print(f"batched members N={ns} scar blobs={[m['n_scar'] for m in members]}", flush=True)
    assert max(ns) > 400_000
    packed = _run_chain(torch, members)
    for m, (lat_p, egm_p) in zip(members, packed):
        lat_a, egm_a = _run_chain(torch, [m])[0]
        assert np.array_equal(lat_a, lat_p)
        assert np.array_equal(egm_a, egm_p)
        assert np.all(np.isfinite(egm_p))
    big = int(np.argmax(ns))
    plain = runtime.PlainSim(members[big]["prep"], tol=1e-15, samples=False)
    plain.run()
    n = _lat_within_one(packed[big][0], plain.lat.cpu().numpy(), members[big]["prep"].n_steps)
    print(f"largest member: LAT within 1 step of plain FP64 over {n} activations")
    assert n > 0.5 * ns[big]


@pytest.mark.timeout(600)
def test_served_requests_from_the_served_ranges(cuda):
    from pyheartsim_amd import runtime, serve
    params = serve.request_params(6, qps=50.0, seed=7, n_range=(2e4, 2e6),
                                  duration_range=(40.0, 80.0), electrode_range=(32, 256))
    reqs = [serve.build_request(p) for p in params]
    ns = [r.prep.n_voxel for r in reqs]
    print(f"served N={ns} models={[r.model for r in reqs]} "
          f"electrodes={[len(r.electrodes) for r in reqs]}", flush=True)
    assert max(ns) > 500_000 and {r.model for r in reqs} == {0, 1}
    srv = serve.Server(max_voxels=sum(ns), max_sims=8, hbm_budget_bytes=100e9, ring=64,
                       chunk_steps=100, keep_results=True)
    st = srv.run(reqs)
    srv.arena.close()
    assert st.completed == len(reqs)
    by_id = {r["rid"]: r for r in st.records}
    for probe in (reqs[int(np.argmax(ns))], next(r for r in reqs if r.model == 1)):
        alone = serve.Server(max_voxels=probe.prep.n_voxel, max_sims=1, hbm_budget_bytes=100e9,
                             ring=64, chunk_steps=100, keep_results=True)
        ra = alone.run([serve.Request(**{**probe.__dict__, "arrival": 0.0})]).records[0]
        alone.arena.close()
        rm = by_id[probe.rid]
        assert np.array_equal(ra["egm"].numpy(), rm["egm"].numpy())
        assert np.array_equal(ra["lat"].numpy(), rm["lat"].numpy())
        plain = runtime.PlainSim(probe.prep, tol=1e-15, samples=False)
        plain.run()
        n = _lat_within_one(rm["lat"].numpy(), plain.lat.cpu().numpy(), probe.prep.n_steps)
        print(f"request {probe.rid} (N={probe.prep.n_voxel}, model {probe.model}): LAT within "
              f"1 step over {n} activations")
        assert n > 0
