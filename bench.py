#!/usr/bin/env python
"""Throughput of the simulation hot path on one MI355X.

Headline workload (default, `--workload batched`): the batched form.  64 seeded
Mitchell-Schaeffer ellipsoidal shells of uneven size (semi-axes U(15, 30) mm, wall U(2, 4) mm,
0.4 mm voxels, 0-3 scar blobs, a seeded pacing site, cycle length U(250, 600) ms), each 1000 ms
at dt = 0.05 ms (20 000 steps) with a 64-electrode electrogram sampled at 1 kHz, FP32
simulation + BF16 hi/lo MFMA electrogram, packed into one arena whose brick shape comes from
the autotuner, steps replayed from hipGraphs.

A step is one dt of reaction + the full CN solve for every voxel of every resident
simulation, plus the electrogram contraction of the samples that step produces (contracted
after every graph chunk from a per-simulation sample ring).  `value` = voxel-steps/s over
exactly --steps timed steps, inputs resident in HBM.  The same run then

  * finishes all 64 simulations (the sustained rate over those steps is reported too) and
    validates two of them against the plain FP64 form over the full 1000 ms: LAT error in
    steps, and electrogram rel-L2 of the chosen chain (FP32 simulation + BF16 hi/lo MFMA)
    against the plain chain (FP64 simulation + FP32 MFMA);
  * times the plain FP64 form on those simulations (the denominator of the speed-up);
  * times the 0.4 mm shell of 199 446 voxels alone and 16 copies packed;
  * serves a Poisson request stream from the served workload's ranges (p50/p99);
  * times the CPU baseline: the C/OpenMP FP64 restatement of legacy/compute_simulation_cpu.py
    (A rebuilt every step, BiCGSTAB with the reference's stop rule) on one member shell from
    a mid-activation state, on the host cores this job may use.

Other workloads: `shell` (the 0.4 mm shell alone and x16, full 40 000-step validation),
`precision` (0.2 mm shell with scar, FP64-residual refinement, 256-electrode electrogram),
`egm` (electrogram alone), `serve` (a longer served stream).  One JSON line goes to stdout.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
BF16_PEAK_TFLOPS = 2500.0
# sizeof(BrickRec): masks, gather bases, slot and temporary-block lanes of one 64-cell brick
BRICK_RECORD_BYTES = 192
METRIC = "voxel-steps/s at LAT err<=1dt (chosen vs plain FP64); p50/p99 sim latency"

PROFILES = os.path.join(ROOT, "profiles")
PROFILE_TRAFFIC = (os.path.join(PROFILES, "arena_sweep_pmc_batched64.md"),
                   os.path.join(PROFILES, "arena_sweep_pmc_shell_ms_x16.md"))
PROFILE_TRAFFIC_EGM = (os.path.join(PROFILES, "egm_shell_e64_t2000.md"),
                       os.path.join(PROFILES, "egm_shell_e256_t2048.md"))


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def _profile_record(path, match, key):
    """`key` of the `traffic:` JSON line of a committed rocprofv3 FETCH_SIZE / WRITE_SIZE
    summary (calibrated on a device copy) whose fields equal `match`; else None."""
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("traffic:"):
                    rec = json.loads(line[len("traffic:"):])
                    if all(rec.get(k) == v for k, v in match.items()):
                        return rec[key]
    except (OSError, ValueError, KeyError):
        pass
    return None


def profiled_traffic(cnt, kind="mid_sweep"):
    """HBM bytes (read + write) of one launch of kernel `kind` (rhs, first_sweep, mid_sweep,
    last_sweep), when a committed summary was measured on an arena of the same voxels and
    bricks."""
    match = {"voxels": cnt["voxels"], "bricks": cnt["bricks"], "record_bytes": BRICK_RECORD_BYTES}
    for path in PROFILE_TRAFFIC:
        per = _profile_record(path, match, "kernels")
        if per is not None and per.get(kind) is not None:
            return per[kind]
        if kind == "mid_sweep":
            r = _profile_record(path, match, "sweep_bytes_per_launch")
            if r is not None:
                return r
    return None


def egm_traffic(n_voxel, electrodes, samples):
    """HBM bytes of one electrogram launch, when a committed summary was measured on the same
    contraction shape."""
    for path in PROFILE_TRAFFIC_EGM:
        r = _profile_record(path, {"n_voxel": n_voxel, "electrodes": electrodes,
                                   "samples": samples}, "bytes_per_launch")
        if r is not None:
            return r
    return None


def host_threads() -> int:
    """Threads this job may run on the host: OMP_NUM_THREADS when set (the share of the host
    given to the job), else every core in the process's affinity mask."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        return int(env)
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def lat_error(a, b, n_steps=None):
    """LAT difference (steps) of two (beats x voxels) tables, -1 = no activation.  With
    n_steps, an activation in the last step of the window that the other form lacks counts
    as a 1-step difference (the other form crosses one step later, outside the window)."""
    a = np.asarray(a).astype(np.int64)
    b = np.asarray(b).astype(np.int64)
    both = (a >= 0) & (b >= 0)
    only = (a >= 0) != (b >= 0)
    at_end = 0
    if n_steps is not None:
        last = only & (np.maximum(a, b) == n_steps - 1)
        at_end = int(np.count_nonzero(last))
        only &= ~last
    mismatch = int(np.count_nonzero(only))
    d = np.abs(a[both] - b[both])
    return dict(max=max(int(d.max()) if d.size else 0, 1 if at_end else 0),
                p99=float(np.percentile(d, 99)) if d.size else 0.0,
                activation_mismatch=mismatch, activation_at_window_end=at_end,
                compared=int(d.size))


def sweep_bytes(k, m, guess=2):
    """Algorithmic bytes per voxel of Chebyshev sweep k (1..m-1) of an m-sweep solve: it reads
    d_k (r for the first sweep, k = 1), r (k >= 2) and d_k-1 (k >= 3; at k = 2 it is r / theta);
    a middle sweep writes d_k+1; the last one (k = m - 1) also reads u*, h, the flag byte and
    the kept corrections (d_1 with guess >= 1, the warm start with guess 2) and writes u*, h
    and those corrections (the next step's reaction included)."""
    rd = 4 + (4 if k >= 2 else 0) + (4 if k >= 3 else 0)
    if k < m - 1:
        return rd + 4
    kept = (4 if guess >= 1 else 0) + (4 if guess == 2 else 0)
    return rd + 4 + 4 + 1 + kept + 4 + 4 + kept


def step_bytes(cnt, sweeps, guess=2, samples_per_step=0.0):
    """Algorithmic HBM bytes of one arena step: each array element read or written once by
    each kernel (sweep_bytes), the brick record once per kernel and brick (it carries the scar
    and temporary-block lanes, so no other kernel than the last reads a flag byte), and the
    sample row.  Kernels: right-hand side (u*, warm start in; r out), the first sweep, the
    middle sweeps (their average under "mid") and the last sweep."""
    V, B = cnt["voxels"], cnt["bricks"]
    rec = B * BRICK_RECORD_BYTES
    m = max(int(sweeps), 2)
    mids = [sweep_bytes(k, m, guess) for k in range(2, m - 1)]
    per = dict(rhs=12 * V + rec, first=sweep_bytes(1, m, guess) * V + rec,
               mid=(sum(mids) / len(mids)) * V + rec if mids else None,
               last=sweep_bytes(m - 1, m, guess) * V + rec)
    if m == 2:  # the first sweep is the last
        per["first"] = per["last"]
    total = per["rhs"] + sum(sweep_bytes(k, m, guess) * V + rec for k in range(1, m)) + \
        4 * V * samples_per_step
    return total, per


def step_roofline(cnt, timing, sweeps, ms_per_step, guess=2, isolated=None):
    """Roofline of the dominant kernel of an arena step.  `timing` is [(kind, ms), ...]: the
    launch durations of one step measured in place inside a replayed graph of the timed region
    (HIP events between the launches, on the arena's stream).  The dominant kernel is the kind
    with the largest total time per step (the intermediate Chebyshev sweep, launched m - 3
    times, or the last sweep when that takes longer); achieved = its algorithmic bytes per
    launch / its average launch duration; traffic = its HBM bytes per launch from the committed
    counter summary of the same arena (None when there is none)."""
    total, per = step_bytes(cnt, sweeps, guess)
    key = {"rhs": "rhs", "first_sweep": "first", "mid_sweep": "mid", "last_sweep": "last"}
    kinds = {}
    for kind, ms in timing:
        k = kinds.setdefault(kind, [0, 0.0])
        k[0] += 1
        k[1] += ms
    table = {}
    for kind, (n, ms) in kinds.items():
        avg = ms / n
        b = per.get(key.get(kind, ""), None)
        table[kind] = dict(launches=n, avg_ms=avg, bytes_per_launch=b,
                           gbs=(b / (avg * 1e-3) / 1e9) if b else None,
                           frac=(b / (avg * 1e-3) / 1e9 / HBM_PEAK_GBS) if b else None)
    for kind in table:
        table[kind]["traffic"] = profiled_traffic(cnt, kind)
    dom = max(kinds, key=lambda k: kinds[k][1])
    d = table[dom]
    in_step = sum(ms for _, ms in timing)
    return {
        "bound": "hbm", "kernel": f"{dom} (k_sweep / k_rhs of the bricked FP32 arena)",
        "achieved": d["gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": d["frac"],
        "traffic": d["traffic"],
        "bytes_per_launch": d["bytes_per_launch"], "avg_launch_ms": d["avg_ms"],
        "launches_per_step": d["launches"],
        "share_of_step": kinds[dom][1] / in_step if in_step > 0 else None,
        "timing": "HIP events around each launch of the last step of the timed region's "
                  "last graph replay",
        "kernels": table,
        "step_kernels_ms": in_step,
        "step_algorithmic_bytes": total,
        "step_hbm_frac": total / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS,
        "isolated_kernels_ms": isolated,
    }


# ------------------------------------------------------------------------- workload builders
def copy_probe(torch, mib=1024, reps=5):
    """Device-to-device copy rate (read + write bytes per second, GB/s) of a `mib` MiB buffer,
    timed with events on torch's stream: a practical HBM ceiling beside the nominal peak."""
    n = mib * (1 << 20) // 4
    src = torch.empty(n, dtype=torch.float32, device="cuda").fill_(1.0)
    dst = torch.empty_like(src)
    dst.copy_(src)
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(reps):
        dst.copy_(src)
    b.record()
    b.synchronize()
    ms = a.elapsed_time(b) / reps
    del src, dst
    return dict(gbs=2.0 * n * 4 / (ms * 1e-3) / 1e9, mib=mib,
                method="torch copy_ of a device buffer to another, HIP events, read + write bytes")


def _batched_member(kw):
    """One member of the batched workload, built on the host (runs in a worker process)."""
    from pyheartsim_amd import configs, runtime
    w = configs.shell_ms(**kw)
    g = w["geometry"]
    P = configs.build_P_2d(g, w["sim"], w["hm"])
    prep = runtime.prepare(P.shape[0], P, g, w["sim"], w["arr"])
    return dict(prep=prep, voxel=g["voxel"], c_voxel=w["hm"]["c_voxel"], delta=g["Delta"],
                electrodes=w["electrodes"], n_scar=kw["n_scar"],
                scar_voxels=int(np.count_nonzero(w["arr"]["node_flag"] == 4)),
                cycle_length=kw["cycle_length"])


def _shell_member(seed):
    """The 0.4 mm shell (25/20/18 mm, 3 mm wall, 4 vein holes) paced at seeded site `seed`."""
    from pyheartsim_amd import configs, geometry as geo, runtime
    g = geo.shell((25.0, 20.0, 18.0), 3.0, 0.4, 4, 6.0, 0)
    w = configs.shell_ms(seed=seed, geometry=g)
    P = configs.build_P_2d(g, w["sim"], w["hm"])
    prep = runtime.prepare(P.shape[0], P, g, w["sim"], w["arr"])
    return dict(prep=prep, voxel=g["voxel"], c_voxel=w["hm"]["c_voxel"], delta=g["Delta"],
                electrodes=w["electrodes"])


def build_parallel(fn, items, workers):
    from pyheartsim_amd.serve import parallel_map
    return parallel_map(fn, items, workers)


# ------------------------------------------------------------------------- chosen-form chain
class Chain:
    """An arena of simulations with their electrograms: steps in hipGraph chunks and, after
    each chunk, contracts the sample rows the chunk produced (read from each simulation's
    sample ring) into that simulation's electrogram on the same stream."""

    def __init__(self, torch, runtime, members, solver, ring=32, egm=True):
        from pyheartsim_amd.simulation.unipolar_electrogram_gpu import LeadField, StreamContraction
        self.torch = torch
        self.members = members
        self.ring = ring
        n_vox = sum(m["prep"].n_voxel for m in members)
        self.solver = solver
        self.arena = runtime.Arena(n_vox, len(members), solver)
        self.slots = [self.arena.add(m["prep"], samples=egm, ring=ring if egm else 0)
                      for m in members]
        self.lfs, self.egms, self.rows = [], [], []
        self.stream = StreamContraction()
        if egm:
            for m, s in zip(members, self.slots):
                nb6 = m["prep"].nb
                self.lfs.append(LeadField(m["electrodes"], m["voxel"], None, m["c_voxel"],
                                          m["delta"], nb6, perm=self.arena.slots[s].perm))
                self.egms.append(torch.empty((m["prep"].n_saved(), len(m["electrodes"])),
                                             dtype=torch.float32, device="cuda"))
                self.rows.append(0)
        self.egm = egm
        self.steps = 0   # steps launched (every member starts at step 0)

    def produced(self, i):
        p = self.members[i]["prep"]
        done = min(p.n_steps, self.steps)
        nsp = p.steps_per_sample
        return min(p.n_saved(), (done + nsp - 1) // nsp if nsp > 0 else done)

    def _jobs(self, upto_of):
        jobs = []
        for i, lf in enumerate(self.lfs):
            upto = upto_of(i)
            su = self.arena.slots[self.slots[i]].su
            r = self.rows[i]
            while r < upto:
                r0 = r % self.ring
                n = min(upto - r, self.ring - r0, 16)
                jobs.append((lf, su[r0:r0 + n], self.egms[i][r:r + n]))
                r += n
        return jobs

    def contract(self):
        """Electrogram rows of every sample produced since the last call: one streamed
        contraction over all simulations (phs_egm_stream)."""
        jobs = self._jobs(self.produced)
        if jobs:
            self.stream.run(jobs)
        for i in range(len(self.lfs)):
            self.rows[i] = max(self.rows[i], self.produced(i))

    def advance(self, n, chunk):
        """n steps in graph chunks of `chunk` (the remainder as eager steps), contracting
        after each chunk."""
        done = 0
        while done < n:
            c = min(chunk, n - done)
            self.arena.step(c, graph=(c == chunk))
            self.steps += c
            done += c
            if self.egm:
                self.contract()

    def advance_marked(self):
        """One step with HIP events around each launch, then the contraction of its rows."""
        self.arena.step_marked()
        self.steps += 1
        if self.egm:
            self.contract()

    def warm_egm(self, rows):
        """Size the streamed contraction's workspace for chunks of up to `rows` rows per
        member (and a ring wrap), so that no timed region allocates."""
        saved = list(self.rows)
        self.rows = [0] * len(self.lfs)
        jobs = self._jobs(lambda i: min(rows, self.egms[i].shape[0]))
        self.stream.reserve(jobs + jobs)
        self.rows = saved

    def normalized_egm(self, i):
        from pyheartsim_amd import _lib
        e = self.egms[i].clone()
        _lib.check(self.arena.lib.phs_egm_normalize(_lib.ptr(e), e.shape[0], e.shape[1], 1.0,
                                                    _lib.stream_ptr()), "phs_egm_normalize")
        return e

    def close(self):
        for lf in self.lfs:
            lf.close()
        self.arena.close()


def timed_chain(torch, chain, steps, warmup, chunk):
    """W untimed warm-up steps, then exactly `steps` steps timed between device
    synchronisations: steps - 1 replayed from graphs of up to `chunk` steps (captured
    beforehand) and the last one launched eagerly with HIP events around each of its launches
    (read afterwards with arena.step_timing())."""
    if steps <= 0:
        raise SystemExit("--steps must be > 0")
    c = max(1, min(chunk, steps - 1))
    chain.arena.prepare_graph(c)
    if warmup > 0:
        chain.advance(warmup, min(chunk, warmup))
        chain.arena.prepare_graph(c)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    if steps > 1:
        chain.advance(steps - 1, c)
    chain.advance_marked()
    torch.cuda.synchronize()
    return time.perf_counter() - t0


def cpu_baseline(kw, u0, h0, step0, budget_s):
    """C/OpenMP FP64 restatement of legacy/compute_simulation_cpu.py (A = I - dt/2 L formed
    every step, BiCGSTAB with the reference stop rule) on the host threads this job may use,
    from a mid-activation state of one member shell, for ~budget_s seconds."""
    from oracle import cpu_oracle
    from pyheartsim_amd import configs
    threads = host_threads()
    w = configs.shell_ms(**kw)
    P = configs.build_P_2d(w["geometry"], w["sim"], w["hm"])
    cs = cpu_oracle.CpuSim(P, w["geometry"], w["sim"], w["arr"])
    cs.u[:] = u0
    cs.h[:] = h0
    cs.step = step0
    cs.run(2, threads)                 # warm-up (first touch of the operator arrays)
    chunk, steps, fails = 20, 0, 0
    t0 = time.perf_counter()
    while True:
        fails += cs.run(chunk, threads, rebuild=True)
        steps += chunk
        dt = time.perf_counter() - t0
        if dt >= budget_s or steps >= 20000:
            break
    n = cs.n
    return dict(value=n * steps / dt, unit="voxel-steps/s", cores=threads, kind="port",
                sample=f"{steps} steps from step {step0} (mid-activation) of one {n}-voxel member "
                       "shell of the batched workload: C/OpenMP FP64 restatement of "
                       "legacy/compute_simulation_cpu.py, CSR L, A = I - dt/2 L formed every step, "
                       "BiCGSTAB(atol 1e-6, rtol 1e-5, maxiter 1000) from x0 = u*",
                cpu_model=cpu_model(), os_cpu_count=os.cpu_count(),
                omp_num_threads=os.environ.get("OMP_NUM_THREADS"),
                build="gcc -O3 -march=x86-64-v2 -fopenmp -ffp-contract=off",
                maxiter_hits=int(fails), seconds=dt,
                tight=cpu_tight(w, P, u0, h0, step0, threads, max(budget_s / 5, 0.5)),
                egm=cpu_egm(w, cs.u, max(budget_s / 5, 0.5)))


def cpu_tight(w, P, u0, h0, step0, threads, budget_s):
    """The same restatement with the CN solve converged tightly (BiCGSTAB rtol 1e-12,
    atol 1e-14), the accuracy the device forms are held to, from the same state."""
    from oracle import cpu_oracle
    cs = cpu_oracle.CpuSim(P, w["geometry"], w["sim"], w["arr"], rtol=1e-12, atol=1e-14)
    cs.u[:] = u0
    cs.h[:] = h0
    cs.step = step0
    chunk, steps, fails = 5, 0, 0
    t0 = time.perf_counter()
    while True:
        fails += cs.run(chunk, threads, rebuild=True)
        steps += chunk
        dt = time.perf_counter() - t0
        if dt >= budget_s or steps >= 5000:
            break
    return dict(value=cs.n * steps / dt, unit="voxel-steps/s", steps=steps, seconds=dt,
                maxiter_hits=int(fails),
                sample="BiCGSTAB(atol 1e-14, rtol 1e-12, maxiter 1000), same member and state")


def cpu_egm(w, u, budget_s):
    """The FP64 NumPy restatement of the electrogram (legacy/compute_unipolar_electrogram_cpu.py,
    unipolar_electrogram_equation_parts.py) on the member's 64 electrodes, timed over calls of
    32 sample rows (copies of the state u) for ~budget_s seconds (at least one call, whose
    distances and lead-field terms are formed anew as the reference's function does)."""
    from oracle import egm_oracle
    g = w["geometry"]
    el = np.asarray(w["electrodes"], np.float64)
    n = g["voxel"].shape[0]
    D = np.zeros((9, n))
    D[0] = D[4] = D[8] = 1.0
    c = np.asarray(w["hm"]["c_voxel"], np.float64)
    rows = np.tile(np.asarray(u, np.float64), (32, 1))
    done = 0
    t0 = time.perf_counter()
    while True:
        egm_oracle.electrogram_raw(el, g["voxel"], D, c, rows, g["Delta"], g["neighbor_id_2d"])
        done += rows.shape[0]
        dt = time.perf_counter() - t0
        if dt >= budget_s:
            break
    E = el.shape[0]
    return dict(value=done / dt, unit="electrogram rows/s", electrodes=E, n_voxel=n,
                useful_gflops=2.0 * E * 3 * n * done / dt / 1e9, rows=done, seconds=dt,
                sample="NumPy FP64 (BLAS threads of the job), calls of 32 rows, distances, "
                       "lead-field terms and gradients included")


def sample_error(torch, runtime, solver, prep, plain_su):
    """Vm error of the chosen form over every 1 kHz sample of the whole run: the simulation
    re-run alone (bitwise what it computes packed) with all its samples kept, against the
    plain FP64 form's samples (caller voxel order).  max and p99 of |dVm| over all
    (sample, voxel) pairs, the p99 on a seeded subsample of 4 M pairs."""
    a = runtime.Arena(prep.n_voxel, 1, solver)
    s = a.add(prep, samples=True)
    while a.steps_done(s) < prep.n_steps:
        a.step(100, graph=True)
    su = a.outputs(s)[0]
    d = (su - plain_su).abs()
    n = d.numel()
    g = torch.Generator(device="cuda").manual_seed(5)
    pick = torch.randint(0, n, (min(n, 4_000_000),), device="cuda", generator=g)
    p99 = float(torch.quantile(d.reshape(-1)[pick], 0.99).item())
    out = dict(max=float(d.max().item()), p99=p99,
               rel_max=float(d.max().item() / plain_su.abs().max().item()),
               samples=int(su.shape[0]))
    a.close()
    del su, d
    return out


def validate_member(torch, runtime, chain, i, lf_cls, cpu_state_step=None):
    """Plain FP64 run of member i over its whole duration (timed), compared with the chosen
    form already finished in `chain`: LAT error and the electrogram chain rel-L2.  Returns
    (record, plain rate, (u, h) at cpu_state_step or None)."""
    m = chain.members[i]
    prep = m["prep"]
    plain = runtime.PlainSim(prep, tol=1e-15, samples=True, sample_h=False, sample_fp64=False)
    state = None
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    if cpu_state_step:
        plain.run(cpu_state_step)
        torch.cuda.synchronize()
        t_pause = time.perf_counter()
        u, h = plain.state()
        state = (u.cpu().numpy(), h.cpu().numpy())
        t0 += time.perf_counter() - t_pause
    plain.run()
    torch.cuda.synchronize()
    t_plain = time.perf_counter() - t0
    rate = prep.n_voxel * prep.n_steps / t_plain
    lat_c = chain.arena.lat(chain.slots[i]).cpu().numpy()
    le = lat_error(lat_c, plain.lat.cpu().numpy(), prep.n_steps)
    lf = lf_cls(m["electrodes"], m["voxel"], None, m["c_voxel"], m["delta"], prep.nb)
    e_ref = lf.run(plain.su, precision="fp32")
    e_c = chain.normalized_egm(i)
    torch.cuda.synchronize()
    rel = float(((e_c - e_ref).norm() / e_ref.norm()).item())
    dvm = float((chain.arena.state(chain.slots[i])[0].double() - plain.state()[0]).abs().max()
                .item())
    lf.close()
    samples = sample_error(torch, runtime, chain.solver, prep, plain.su)
    rec = dict(member=i, n_voxel=prep.n_voxel, steps=prep.n_steps, n_scar=m.get("n_scar"),
               scar_voxels=m.get("scar_voxels"), cycle_length_ms=m.get("cycle_length"),
               lat_err_steps=le, egm_rel_l2=rel, max_abs_dVm_final=dvm, dVm_samples=samples,
               plain_voxel_steps_per_s=rate, plain_sweeps=plain.info()[0])
    del plain
    return rec, rate, state


def shell_legs(torch, runtime, solver, workers, steps=400, warmup=100, chunk=50):
    """The 0.4 mm shell (199 446 voxels): plain FP64, chosen form alone, and 16 copies packed
    (different seeded pacing sites), each timed over `steps` steps without electrograms."""
    members = build_parallel(_shell_member, list(range(16)), workers)
    prep = members[0]["prep"]
    N = prep.n_voxel
    out = dict(n_voxel=N)
    plain = runtime.PlainSim(prep, tol=1e-15, samples=False)
    plain.run(warmup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    plain.run(steps)
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    out["plain_fp64"] = dict(voxel_steps_per_s=N * steps / t, ms_per_step=1e3 * t / steps,
                             sweeps=plain.info()[0])
    del plain
    for name, ms in (("alone", members[:1]), ("x16", members)):
        ch = Chain(torch, runtime, ms, solver, egm=False)
        t = timed_chain(torch, ch, steps, warmup, chunk)
        v = ch.arena.counts()["voxels"]
        out[name] = dict(voxel_steps_per_s=v * steps / t, ms_per_step=1e3 * t / steps,
                         sweeps=ch.arena.sweeps(), voxels=v)
        if name == "alone":
            total, _ = step_bytes(ch.arena.counts(), ch.arena.sweeps(), solver.guess)
            out[name]["step_hbm_frac"] = total / (t / steps) / 1e9 / HBM_PEAK_GBS
        ch.close()
        torch.cuda.empty_cache()
    return out


def serve_leg(torch, runtime, solver, reqs, qps, max_voxels=60_000_000):
    from pyheartsim_amd import serve
    srv = serve.Server(max_voxels=min(max_voxels, sum(r.prep.n_voxel for r in reqs)),
                       max_sims=64, hbm_budget_bytes=200e9, solver=solver, ring=256,
                       chunk_steps=100)
    st = srv.run(reqs, progress=lambda r: log(f"  served rid={r['rid']} N={r['n_voxel']} "
                                               f"steps={r['steps']} latency={r['latency']:.2f}s"))
    srv.arena.close()
    ns = [r.prep.n_voxel for r in reqs]
    return dict(qps=qps, requests=st.completed, p50_s=st.p50, p99_s=st.p99,
                voxel_steps_per_s=st.throughput, wall_s=st.wall,
                compaction_moves=dict(bricks=st.brick_moves, voxels=st.voxel_moves),
                n_voxel_range=[min(ns), max(ns)],
                duration_ms_range=[min(r.duration_ms for r in reqs),
                                   max(r.duration_ms for r in reqs)],
                workload="Poisson arrivals; shells of 2e4-2e6 voxels (log-uniform) at 0.4 mm, "
                         "MS/AP 50/50, 500-5000 ms, 32-256 electrodes")


# ------------------------------------------------------------------------- workloads
def run_batched(args, torch, runtime, pre):
    """The headline: 64 uneven shells with electrograms (see the module docstring)."""
    from pyheartsim_amd import autotune
    from pyheartsim_amd.simulation.unipolar_electrogram_gpu import LeadField
    members, kws, reqs = pre["members"], pre["kws"], pre.get("requests")
    ns = [m["prep"].n_voxel for m in members]
    base = runtime.SolverConfig(sweeps=max(args.sweeps, 0), guess=args.guess)
    t0 = time.perf_counter()
    solver, tune_table = autotune.tune([m["prep"] for m in members[::8]], base, steps=200,
                                       warmup=50, graph_steps=50)
    tune = dict(brick=list(solver.brick), bricks_per_wave=solver.bricks_per_wave,
                seconds=time.perf_counter() - t0,
                candidates=[dict(brick=r["brick"], bricks_per_wave=r["bricks_per_wave"],
                                 fill=r["fill"], voxel_steps_per_s=r["voxel_steps_per_s"])
                            for r in tune_table])
    log(f"autotune: {tune}")
    torch.cuda.empty_cache()
    t0 = time.perf_counter()
    chain = Chain(torch, runtime, members, solver, ring=args.ring)
    cnt = chain.arena.counts()
    sweeps = chain.arena.sweeps()
    nspm = members[0]["prep"].steps_per_sample
    G = args.graph_steps
    assert math.ceil(G / nspm) + 1 <= args.ring, "sample ring too short for the graph chunk"
    chain.warm_egm(math.ceil(G / nspm) + 1)
    log(f"arena: {cnt}, sweeps {sweeps}, added in {time.perf_counter()-t0:.1f}s")
    K, W = args.steps, args.warmup
    t = timed_chain(torch, chain, K, W, G)
    value = cnt["voxels"] * K / t
    ms_step = 1e3 * t / K
    log(f"batched: {value:.4g} voxel-steps/s, {ms_step:.4f} ms/step over {K} steps")
    timing = chain.arena.step_timing()
    log(f"step launches (ms): {timing}")
    if args.no_validate:
        prof = chain.arena.profile(50)
        cnt_line = dict(cnt)
        chain.close()
        return {"metric": METRIC, "value": value, "unit": "voxel-steps/s", "n_gpus": 1,
                "steps": K, "warmup": W, "ms_per_step": ms_step, "higher_is_better": True,
                "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
                "config": {"workload": "batched (timed region only, not validated)",
                           "voxels": cnt_line["voxels"], "sweeps": sweeps},
                "roofline": step_roofline(cnt, timing, sweeps, ms_step, args.guess, prof),
                "cpu_baseline": None}
    # sustained rate: the rest of the 1000 ms of every member, electrograms included
    n_total = members[0]["prep"].n_steps
    rest = n_total - chain.steps
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    chain.advance(rest, G)
    torch.cuda.synchronize()
    t_rest = time.perf_counter() - t0
    sustained = dict(voxel_steps_per_s=cnt["voxels"] * rest / t_rest,
                     ms_per_step=1e3 * t_rest / rest, steps=rest)
    log(f"sustained: {sustained}")
    prof = chain.arena.profile(50)
    roof = step_roofline(cnt, timing, sweeps, ms_step, args.guess, prof)
    roof["sustained_step_hbm_frac"] = (roof["step_algorithmic_bytes"] /
                                       (sustained["ms_per_step"] * 1e-3) / 1e9 / HBM_PEAK_GBS)
    cp = copy_probe(torch)
    roof["copy_probe"] = cp
    roof["frac_of_copy"] = roof["achieved"] / cp["gbs"] if roof["achieved"] else None
    log(f"kernels: {prof}; copy probe {cp['gbs']:.0f} GB/s")
    # validation against the plain FP64 form: the largest member and the most scarred one
    big = int(np.argmax(ns))
    scarred = max((i for i in range(len(members)) if i != big),
                  key=lambda i: members[i]["scar_voxels"])
    val, plain_rates, cpu_state = [], [], None
    for i in (big, scarred):
        rec, rate, st = validate_member(torch, runtime, chain, i, LeadField,
                                        cpu_state_step=args.cpu_from_step if i == big else None)
        val.append(rec)
        plain_rates.append(rate)
        if st is not None:
            cpu_state = st
        log(f"validate: {rec}")
    egm_flop = sum(2.0 * len(m["electrodes"]) * m["prep"].n_saved() * m["prep"].n_voxel
                   for m in members)
    chain.close()
    del chain
    torch.cuda.empty_cache()
    plain_rate = sum(v["plain_voxel_steps_per_s"] * v["n_voxel"] for v in val) / \
        sum(v["n_voxel"] for v in val)

    shell = None
    if not args.only_headline:
        shell = shell_legs(torch, runtime, solver, args.workers)
        log(f"shell: {shell}")
    serve_res = None
    if reqs:
        serve_res = serve_leg(torch, runtime, solver, reqs, args.qps)
        log(f"serve: {serve_res}")
    cpu = None
    if args.cpu_seconds > 0 and cpu_state is not None:
        cpu = cpu_baseline(kws[big], cpu_state[0], cpu_state[1], args.cpu_from_step,
                           args.cpu_seconds)
        log(f"cpu: {cpu}")
    return {
        "metric": METRIC,
        "value": value,
        "unit": "voxel-steps/s",
        "n_gpus": 1, "steps": K, "warmup": W, "ms_per_step": ms_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"batched: {len(members)} seeded MS ellipsoidal shells (semi-axes "
                        f"15-30 mm, wall 2-4 mm, 0.4 mm voxels, 0-3 scar blobs, CL 250-600 ms), "
                        f"N {min(ns)}-{max(ns)} (total {sum(ns)}), 1000 ms at dt 0.05 ms, "
                        f"64-electrode BF16 hi/lo electrograms at 1 kHz, one arena, hipGraph steps",
            "sims": len(members), "voxels": sum(ns), "n_voxel_range": [min(ns), max(ns)],
            "steps_per_sim": n_total, "sweeps": sweeps, "guess": args.guess,
            "brick": list(solver.brick), "bricks_per_wave": solver.bricks_per_wave,
            "graph_steps": G, "electrodes": 64, "egm_precision": "bf16x3",
            "sample_ring_rows": args.ring,
        },
        "speedup_vs_plain_fp64": value / plain_rate,
        "plain_fp64": dict(voxel_steps_per_s=plain_rate, members=[v["member"] for v in val]),
        "sustained": sustained,
        "validate": val,
        "egm_useful_tflop_total": egm_flop / 1e12,
        "autotune": tune,
        "shell": shell,
        "latency_ms": ({"p50": 1e3 * serve_res["p50_s"], "p99": 1e3 * serve_res["p99_s"]}
                       if serve_res else {"p50": None, "p99": None}),
        "serve": serve_res,
        "roofline": roof,
        "cpu_baseline": cpu,
    }


def run_shell(args, torch, runtime, pre):
    """The 0.4 mm shell alone (plain, chosen single, x16) with the full 40 000-step validation
    of the chosen form against plain FP64 and the 64-electrode electrogram chain."""
    from pyheartsim_amd.simulation.unipolar_electrogram_gpu import LeadField
    solver = runtime.SolverConfig(sweeps=max(args.sweeps, 0), guess=args.guess)
    members = pre["members"]
    chain = Chain(torch, runtime, members[:1], solver, ring=args.ring)
    nspm = members[0]["prep"].steps_per_sample
    chain.warm_egm(math.ceil(args.graph_steps / nspm) + 1)
    t = timed_chain(torch, chain, args.steps, args.warmup, args.graph_steps)
    timing = chain.arena.step_timing()
    N = members[0]["prep"].n_voxel
    value = N * args.steps / t
    chain.advance(members[0]["prep"].n_steps - chain.steps, args.graph_steps)
    rec, rate, _ = validate_member(torch, runtime, chain, 0, LeadField)
    cnt = chain.arena.counts()
    prof = chain.arena.profile(50)
    sweeps = chain.arena.sweeps()
    chain.close()
    legs = shell_legs(torch, runtime, solver, args.workers)
    return {
        "metric": METRIC, "value": value, "unit": "voxel-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1e3 * t / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"shell: MS ellipsoidal shell 25/20/18 mm, 3 mm wall, 4 vein "
                               f"holes, 0.4 mm voxels (N={N}), dt 0.05 ms, CL 500 ms, 2000 ms, "
                               f"64-electrode electrogram, alone", "n_voxel": N,
                   "sweeps": sweeps, "guess": args.guess},
        "speedup_vs_plain_fp64": value / rate, "validate": rec, "legs": legs,
        "roofline": step_roofline(cnt, timing, sweeps, 1e3 * t / args.steps, args.guess, prof),
        "cpu_baseline": None,
    }


def run_egm(args, torch, runtime, pre):
    """The electrogram alone: E electrodes x T samples over the 0.4 mm shell's 199 446 voxels or
    the 0.2 mm shell's 1.6 M (BF16 hi/lo MFMA contraction with the precomputed lead field), synthetic
    plane-wave samples, --steps timed launches after --warmup; rel-L2 against the FP32 MFMA
    form on the same samples.  One launch reads the samples (4 T N bytes) and the BF16 hi/lo
    lead field (4 E N) and writes the T x E result."""
    from pyheartsim_amd import geometry as geo
    from pyheartsim_amd.simulation.unipolar_electrogram_gpu import LeadField
    E, T = args.electrodes, args.samples
    el = geo.fibonacci_electrodes(E, 15.0)
    if args.egm_geometry == "fine":
        g = pre["fine"]["geometry"]
        voxel, name = g["voxel"], "0.2 mm precision shell"
        lf = LeadField(el, voxel, None, pre["fine"]["hm"]["c_voxel"], g["Delta"],
                       g["neighbor_id_2d"])
    else:
        m = pre["members"][0]
        voxel, name = m["voxel"], "0.4 mm shell"
        lf = LeadField(el, voxel, None, m["c_voxel"], m["delta"], m["prep"].nb)
    N = voxel.shape[0]
    x = torch.from_numpy(voxel[:, 0].astype(np.float32)).cuda()
    tt = torch.arange(T, device="cuda", dtype=torch.float32)[:, None]
    samples = torch.sigmoid(tt * 0.03 - (x[None, :] + 25.0) * 0.3)   # a plane wave in x
    e_ref = lf.run(samples, precision="fp32", normalize=False)
    out = torch.empty((T, E), dtype=torch.float32, device="cuda")
    for _ in range(max(args.warmup, 1)):
        lf.run(samples, precision="bf16x3", normalize=False, out=out)
    K = max(args.steps, 1)
    st = torch.cuda.current_stream()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    ev0.record(st)
    for _ in range(K):
        lf.run(samples, precision="bf16x3", normalize=False, out=out)
    ev1.record(st)
    torch.cuda.synchronize()
    ms = ev0.elapsed_time(ev1) / K
    rel = float(((out - e_ref).norm() / e_ref.norm()).item())
    useful = 2.0 * E * T * N
    nbytes = 4.0 * T * N + 4.0 * E * N + 4.0 * T * E
    gbs = nbytes / (ms * 1e-3) / 1e9
    issued = 3.0 * useful / (ms * 1e-3) / 1e12
    mfma_bound = E >= 128
    log(f"egm: E={E} T={T} N={N} {ms:.4f} ms/launch, rel_l2 {rel:.3e}, {gbs:.0f} GB/s, "
        f"{issued:.0f} TF/s issued")
    roof = ({"bound": "mfma", "achieved": issued, "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
             "frac": issued / BF16_PEAK_TFLOPS} if mfma_bound else
            {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
             "frac": gbs / HBM_PEAK_GBS})
    roof.update(kernel="egm_gemm_lds + egm_reduce (one electrogram)", bytes_per_launch=nbytes,
                hbm_gbs=gbs, mfma_issued_tflops=issued,
                traffic=egm_traffic(N, E, T))
    return {
        "metric": "electrogram useful TFLOP/s (2 E T N per launch, BF16 hi/lo MFMA)",
        "value": useful / (ms * 1e-3) / 1e12, "unit": "TFLOP/s",
        "n_gpus": 1, "steps": K, "warmup": max(args.warmup, 1), "ms_per_step": ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "bf16x3",
        "data": "synthetic",
        "config": {"workload": f"egm: {name} lead field, E={E} electrodes, T={T} samples, "
                               f"N={N} voxels", "electrodes": E, "samples": T, "n_voxel": N},
        "roofline": roof, "rel_l2_vs_fp32_mfma": rel, "cpu_baseline": None,
    }


def run_precision(args, torch, runtime, pre):
    """The 0.2 mm shell with 3 scar blobs (15 % of the wall): FP32 sweeps without and with
    FP64-residual refinement, the achieved CN residual, LAT against the plain FP64 form over
    the window, and the 256-electrode electrogram of the simulated samples (chosen chain)
    against the plain chain."""
    from pyheartsim_amd import configs
    from pyheartsim_amd.simulation.unipolar_electrogram_gpu import LeadField
    w = configs.shell_fine(t_final=args.precision_ms)
    g = w["geometry"]
    P = configs.build_P_2d(g, w["sim"], w["hm"])
    prep = runtime.prepare(P.shape[0], P, g, w["sim"], w["arr"])
    N = prep.n_voxel
    out = dict(workload="precision", n_voxel=N, steps=prep.n_steps)
    plain = runtime.PlainSim(prep, tol=1e-15, samples=True, sample_h=False, sample_fp64=False)
    t0 = time.perf_counter()
    plain.run()
    torch.cuda.synchronize()
    plain_rate = N * prep.n_steps / (time.perf_counter() - t0)
    out["plain"] = dict(voxel_steps_per_s=plain_rate, sweeps=plain.info()[0],
                        lambda_max=plain.info()[2])
    ref = plain.lat.cpu().numpy()
    u_plain = plain.state()[0]
    el = w["electrodes"]
    lf_p = LeadField(el, g["voxel"], None, w["hm"]["c_voxel"], g["Delta"], prep.nb)
    e_plain = lf_p.run(plain.su, precision="fp32")
    lf_p.close()
    del plain
    torch.cuda.empty_cache()
    value = ms_step = None
    for refine, rsw in ((0, 8), (3, 12)):
        cfg = runtime.SolverConfig(sweeps=max(args.sweeps, 0), guess=args.guess, refine=refine,
                                   refine_sweeps=rsw, diag=True)
        members = [dict(prep=prep, voxel=g["voxel"], c_voxel=w["hm"]["c_voxel"],
                        delta=g["Delta"], electrodes=el)]
        chain = Chain(torch, runtime, members, cfg, ring=args.ring)
        lmax = chain.arena.bounds(chain.slots[0])[1]
        chain.warm_egm(math.ceil(args.graph_steps / prep.steps_per_sample) + 1)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        chain.advance(prep.n_steps, args.graph_steps)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        res = chain.arena.last_residual(chain.slots[0])
        lat = chain.arena.lat(chain.slots[0]).cpu().numpy()
        dvm = float((chain.arena.state(chain.slots[0])[0].double() - u_plain).abs().max().item())
        e_c = chain.normalized_egm(0)
        rel = float(((e_c - e_plain).norm() / e_plain.norm()).item())
        rate = N * prep.n_steps / dt
        out[f"refine{refine}"] = dict(
            voxel_steps_per_s=rate, speedup_vs_plain=rate / plain_rate,
            sweeps=chain.arena.sweeps(), refine_sweeps=rsw if refine else 0,
            lambda_max=lmax, cn_rel_residual=res, max_abs_dVm_final=dvm,
            lat_err_steps=lat_error(lat, ref, prep.n_steps), egm_rel_l2_vs_plain_chain=rel)
        log(f"precision refine={refine}: {out[f'refine{refine}']}")
        if refine:
            value, ms_step = rate, 1e3 * dt / prep.n_steps
        chain.close()
        torch.cuda.empty_cache()
    return {"metric": METRIC, "value": value, "unit": "voxel-steps/s", "n_gpus": 1,
            "steps": prep.n_steps, "warmup": 0, "ms_per_step": ms_step,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
            "data": "synthetic",
            "config": {"workload": f"precision: 0.2 mm shell (N={N}) with 3 scar blobs, "
                                   f"{args.precision_ms} ms, refine 3 x 12, 256 electrodes"},
            "detail": out, "cpu_baseline": None, "roofline": None}


def run_serve(args, torch, runtime, pre):
    solver = runtime.SolverConfig(sweeps=max(args.sweeps, 0), guess=args.guess)
    res = serve_leg(torch, runtime, solver, pre["requests"], args.qps)
    return {"metric": METRIC, "value": res["voxel_steps_per_s"], "unit": "voxel-steps/s",
            "n_gpus": 1, "steps": None, "warmup": None, "ms_per_step": None,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
            "data": "synthetic",
            "config": {"workload": f"served: {len(pre['requests'])} requests at QPS {args.qps}"},
            "latency_ms": {"p50": 1e3 * res["p50_s"], "p99": 1e3 * res["p99_s"]},
            "serve": res, "cpu_baseline": None, "roofline": None}


def parser():
    ap = argparse.ArgumentParser(description=__doc__,
                                 formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--workload", choices=["batched", "shell", "precision", "egm", "serve"],
                    default="batched")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--sims", type=int, default=64, help="members of the batched workload")
    ap.add_argument("--sweeps", type=int, default=0,
                    help="Chebyshev sweeps per CN solve; 0 = the arena's automatic count")
    ap.add_argument("--guess", type=int, default=2,
                    help="warm start of the CN correction: 0 none, 1 previous step, 2 linear "
                         "extrapolation of the previous two")
    ap.add_argument("--graph-steps", type=int, default=100,
                    help="steps per replayed hipGraph (and per electrogram contraction)")
    ap.add_argument("--ring", type=int, default=16, help="sample ring rows per simulation")
    ap.add_argument("--serve-requests", type=int, default=64)
    ap.add_argument("--qps", type=float, default=4.0)
    ap.add_argument("--electrodes", type=int, default=64, help="egm workload")
    ap.add_argument("--samples", type=int, default=2000, help="egm workload: time samples")
    ap.add_argument("--egm-geometry", choices=["shell", "fine"], default="shell",
                    help="egm workload: the 0.4 mm shell (199 446 voxels) or the 0.2 mm "
                         "precision shell (1.6 M voxels)")
    ap.add_argument("--precision-ms", type=float, default=300.0)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--cpu-from-step", type=int, default=3000)
    ap.add_argument("--workers", type=int, default=0,
                    help="host processes for building the workloads (0 = the host threads "
                         "this job may use, at most 16)")
    ap.add_argument("--no-validate", action="store_true",
                    help="batched workload: stop after the timed region (profiler runs); the "
                         "line then carries no validation, serve or CPU figures")
    ap.add_argument("--only-headline", action="store_true",
                    help="batched workload only: skip the shell legs")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MB",
                    help="run a device copy of this size first (HBM counter calibration)")
    return ap


def run(argv=None) -> dict:
    """Parse `argv` (default sys.argv), run the workload and return the result record
    (main() prints it as one JSON line)."""
    args = parser().parse_args(argv)
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the simulations "
              "shard by request across GPUs, see the multi-GPU design notes.", file=sys.stderr)
        sys.exit(2)
    if args.steps <= 0 or args.warmup < 0:
        print("--steps must be > 0 and --warmup >= 0", file=sys.stderr)
        sys.exit(2)
    if args.workers <= 0:
        args.workers = min(16, host_threads())

    # host-side workload construction first, in forked workers, before the GPU is touched
    from pyheartsim_amd import configs
    t0 = time.perf_counter()
    pre = {}
    if args.workload == "batched":
        pre["kws"] = configs.ragged_params(args.sims)
        pre["members"] = build_parallel(_batched_member, pre["kws"], args.workers)
    elif args.workload == "egm" and args.egm_geometry == "fine":
        pre["fine"] = configs.shell_fine(t_final=1.0)
    elif args.workload in ("shell", "egm"):
        pre["members"] = build_parallel(_shell_member, [0], 1)
    if args.workload in ("batched", "serve") and args.serve_requests > 0:
        from pyheartsim_amd import serve
        pre["requests"] = serve.make_requests(args.serve_requests, args.qps, seed=11,
                                              workers=args.workers)
    log(f"host setup: {time.perf_counter()-t0:.1f}s ({args.workers} workers)")

    import torch
    from pyheartsim_amd import runtime
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    if args.calibrate_copy > 0:
        n = args.calibrate_copy * (1 << 20) // 4
        src = torch.rand(n, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log(f"calibration copy: {args.calibrate_copy} MiB read + {args.calibrate_copy} MiB written")
        del src, dst
    fn = {"batched": run_batched, "shell": run_shell, "precision": run_precision,
          "egm": run_egm, "serve": run_serve}[args.workload]
    rec = fn(args, torch, runtime, pre)
    rec["device"] = device_info(torch)
    return rec


def device_info(torch) -> dict:
    """The GPU the line was measured on, as the runtime reports it (name, architecture,
    compute units, memory), and the HIP and PyTorch versions."""
    p = torch.cuda.get_device_properties(0)
    return dict(name=p.name, arch=getattr(p, "gcnArchName", None),
                compute_units=p.multi_processor_count, memory_gib=p.total_memory / 2**30,
                l2_cache_mib=getattr(p, "L2_cache_size", 0) / 2**20 or None,
                hip=torch.version.hip, torch=torch.__version__)


def main():
    print(json.dumps(run()), flush=True)


if __name__ == "__main__":
    main()
