"""Continuous-batching request loop over the bricked FP32 arena.

Requests (each a focal-site simulation with its own geometry, model, duration and
electrode set) arrive open-loop at Poisson times.  The loop

  * admits queued requests first-come-first-served whenever the arena has room for
    their voxels and bricks and the HBM budget has room for their footprint
    (state, sample ring, lead field, activation times, electrogram); when the free
    voxels suffice only in total, it first compacts the arena's voxel ranges;
  * keeps every admitted simulation's state on the device and advances all of them
    together in hipGraph-replayed chunks of steps (each simulation on its own step
    counter, so a request admitted mid-run starts at its step 0);
  * streams each simulation's 1 kHz samples through a two-half sample ring into the
    MFMA electrogram (one contraction per half ring), so no request ever holds its
    full (samples x voxels) history;
  * retires a finished simulation: last partial electrogram chunk, normalisation,
    activation times in caller order, results copied to the host; its arena ranges
    are reused by the next admission, and the resident simulations' brick records slide
    down to close the gap (phs_arena_compact), so that the launches shrink with the
    resident set.

Latency of a request = completion time (its results are on the host) - arrival time.
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass, field

import numpy as np

from . import configs, geometry as geo, runtime
from .simulation.unipolar_electrogram_gpu import LeadField


@dataclass
class Request:
    rid: int
    arrival: float          # seconds after the start of the stream
    prep: runtime.Prepared
    voxel: np.ndarray
    neighbor_id_2d: np.ndarray
    delta: float
    c_voxel: np.ndarray
    electrodes: np.ndarray
    model: int
    duration_ms: float

    @property
    def voxel_steps(self) -> int:
        return self.prep.n_voxel * self.prep.n_steps

    def footprint_bytes(self, ring: int) -> int:
        n, e, t = self.prep.n_voxel, len(self.electrodes), self.prep.n_saved()
        kpad = (n + 63) // 64 * 64
        epad = (e + 63) // 64 * 64
        return (n * 33            # arena state per voxel (u, h, r, 2 sweep buffers, warm start, flags, alpha)
                + n * 192 // 40   # brick records (~40 voxels per brick)
                + 4 * ring * n    # sample ring
                + 8 * epad * kpad  # lead field (fp32 + bf16 hi/lo)
                + 28 * n          # activation times (7 beats)
                + 4 * t * e * 2)  # electrogram and split-K workspace


def shell_for_voxels(n_target: float, seed: int, delta: float = 0.4):
    """Seeded ellipsoidal shell whose voxel count is close to n_target (scaled axes)."""
    rng = np.random.default_rng(seed)
    base = np.array([25.0, 20.0, 18.0]) * rng.uniform(0.9, 1.1, size=3)
    s = math.sqrt(n_target / 199446.0) * (0.4 / delta) ** 1.0
    axes = tuple(float(a) for a in base * max(s, 0.3))
    wall = min(3.0, 0.25 * min(axes))
    hole = min(6.0, 0.22 * min(axes))
    return geo.shell(axes, wall, delta, 4, hole, seed)


def request_params(n_requests: int, qps: float, seed: int = 11, n_range=(2e4, 2e6),
                   duration_range=(500.0, 5000.0), electrode_range=(32, 256), delta: float = 0.4,
                   dt_ms: float = 0.05):
    """The draws of the served workload, one dict per request, from one generator in request
    order: Poisson arrival times at `qps`, a log-uniform voxel count, a shell seed,
    Mitchell-Schaeffer or Aliev-Panfilov (50/50), a uniform duration and electrode count."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for rid in range(n_requests):
        t += rng.exponential(1.0 / qps)
        n_target = math.exp(rng.uniform(math.log(n_range[0]), math.log(n_range[1])))
        s = int(rng.integers(1 << 30))
        model = int(rng.integers(2))
        dur = float(rng.uniform(*duration_range))
        n_el = int(rng.integers(electrode_range[0], electrode_range[1] + 1))
        out.append(dict(rid=rid, arrival=t, n_target=n_target, seed=s, model=model,
                        duration_ms=dur, n_electrodes=n_el, delta=delta, dt_ms=dt_ms))
    return out


def build_request(p: dict) -> Request:
    """The Request of one `request_params` entry: its seeded shell, parameters, schedule and
    electrode set (host work only)."""
    g = shell_for_voxels(p["n_target"], p["seed"], p["delta"])
    model = p["model"]
    ts = 6 if model == 1 else 1
    w = configs.shell_ms(delta=p["delta"], t_final=p["duration_ms"], seed=p["seed"],
                         n_electrodes=p["n_electrodes"], dt=p["dt_ms"], heart_model_flag=model,
                         geometry=g, semi_axes=g["semi_axes"], wall=g["wall"])
    P = configs.build_P_2d(g, w["sim"], w["hm"])
    prep = runtime.prepare(P.shape[0], P, g, w["sim"], w["arr"])
    el = geo.fibonacci_electrodes(p["n_electrodes"], 0.6 * min(g["semi_axes"]))
    return Request(p["rid"], p["arrival"], prep, g["voxel"], g["neighbor_id_2d"][:, :6].copy(),
                   g["Delta"], w["hm"]["c_voxel"] / ts, el, model, p["duration_ms"])


def make_requests(n_requests: int, qps: float, seed: int = 11, n_range=(2e4, 2e6),
                  duration_range=(500.0, 5000.0), electrode_range=(32, 256), delta: float = 0.4,
                  dt_ms: float = 0.05, workers: int = 1):
    """The served workload: Poisson arrivals at `qps`; each request a seeded shell of
    log-uniform voxel count, Mitchell-Schaeffer or Aliev-Panfilov (50/50), uniform
    duration and electrode count.  workers > 1 builds the requests in forked processes (call
    it before the process uses the GPU)."""
    params = request_params(n_requests, qps, seed, n_range, duration_range, electrode_range,
                            delta, dt_ms)
    return parallel_map(build_request, params, workers)


def parallel_map(fn, items, workers: int):
    """[fn(x) for x in items], in forked worker processes when workers > 1 (host-only work;
    the parent must not have initialised the GPU)."""
    if workers <= 1 or len(items) <= 1:
        return [fn(x) for x in items]
    import concurrent.futures as cf
    import multiprocessing as mp
    with cf.ProcessPoolExecutor(max_workers=min(workers, len(items)),
                                mp_context=mp.get_context("fork")) as ex:
        return list(ex.map(fn, items))


@dataclass
class _Active:
    req: Request
    slot: int
    lf: LeadField
    egm: object              # device (T x E) electrogram being assembled
    done_rows: int = 0       # samples already contracted
    admitted: float = 0.0
    finished_event: object = None
    results: dict = field(default_factory=dict)


@dataclass
class ServeStats:
    latencies: list
    voxel_steps: int
    wall: float
    admitted: int
    completed: int
    p50: float
    p99: float
    throughput: float
    records: list
    brick_moves: int = 0    # simulations whose brick records compaction moved
    voxel_moves: int = 0    # simulations whose voxel data compaction moved


class Server:
    def __init__(self, max_voxels: int, max_sims: int, hbm_budget_bytes: float,
                 solver: runtime.SolverConfig = runtime.SolverConfig(), ring: int = 256,
                 chunk_steps: int = 100, keep_results: bool = False):
        import torch
        self.torch = torch
        self.arena = runtime.Arena(max_voxels, max_sims, solver, max_bricks=max_voxels // 8)
        self.ring = ring
        self.chunk = chunk_steps
        self.budget = hbm_budget_bytes
        self.in_use = 0
        self.keep = keep_results
        self.brick_moves = 0
        self.voxel_moves = 0

    def _admit(self, r: Request, now: float) -> _Active | None:
        # A chunk of steps must produce at most `ring` samples: the ring holds two halves,
        # one being filled while the other waits for its contraction.
        per_chunk = math.ceil(self.chunk / max(r.prep.steps_per_sample, 1))
        if per_chunk > self.ring:
            raise ValueError(f"request {r.rid}: {per_chunk} samples per chunk of {self.chunk} "
                             f"steps exceed the half ring of {self.ring} rows")
        fp = r.footprint_bytes(2 * self.ring)
        if self.in_use + fp > self.budget:
            return None
        slot = self._add(r)
        if slot is None:
            free_voxels, _, free_slots = self.arena.free_capacity()
            if free_slots == 0 or free_voxels < r.prep.n_voxel:
                return None
            # enough voxels in total, split by the ranges of running simulations: close the
            # gaps (results unchanged) and try once more
            self.voxel_moves += self.arena.compact(0.0, voxels=True)
            slot = self._add(r)
            if slot is None:
                return None
        perm = self.arena.slots[slot].perm
        lf = LeadField(r.electrodes, r.voxel, None, r.c_voxel, r.delta, r.neighbor_id_2d, perm=perm)
        egm = self.torch.empty((r.prep.n_saved(), len(r.electrodes)), dtype=self.torch.float32,
                               device="cuda")
        self.in_use += fp
        return _Active(r, slot, lf, egm, 0, now)

    def _add(self, r: Request):
        """The request's arena slot, or None when the arena has no room (no free slot, or no
        free voxel or brick range large enough)."""
        try:
            return self.arena.add(r.prep, samples=True, ring=2 * self.ring)
        except runtime._lib.PhsError as e:
            if e.code == runtime._lib.PHS_ERR_CAPACITY:
                return None
            raise

    def _contract(self, a: _Active, upto: int):
        """Electrogram rows [done_rows, upto) from the sample ring (upto - done_rows <= ring)."""
        R2 = 2 * self.ring
        su = self.arena.slots[a.slot].su
        while a.done_rows < upto:
            r0 = a.done_rows % R2
            n = min(upto - a.done_rows, R2 - r0)
            a.lf.run(su[r0:r0 + n], precision="bf16x3", normalize=False,
                     out=a.egm[a.done_rows:a.done_rows + n])
            a.done_rows += n

    def run(self, requests, progress=None) -> ServeStats:
        torch = self.torch
        pending = sorted(requests, key=lambda r: r.arrival)
        queue: list[Request] = []
        active: list[_Active] = []
        retiring: list[_Active] = []
        lat = []
        records = []
        voxel_steps = 0
        t_start = time.perf_counter()
        pi = 0
        inflight = []
        while pi < len(pending) or queue or active or retiring:
            now = time.perf_counter() - t_start
            while pi < len(pending) and pending[pi].arrival <= now:
                queue.append(pending[pi])
                pi += 1
            while queue:
                a = self._admit(queue[0], now)
                if a is None:
                    break
                queue.pop(0)
                active.append(a)
            if not active and not retiring:
                if pi < len(pending):
                    time.sleep(max(0.0, min(0.01, pending[pi].arrival - now)))
                continue
            if active:
                self.arena.step(self.chunk, graph=True)
                ev = torch.cuda.Event()
                ev.record()
                inflight.append(ev)
                still = []
                for a in active:
                    done = min(a.req.prep.n_steps, self.arena.steps_done(a.slot))
                    nsp = a.req.prep.steps_per_sample
                    produced = min(a.req.prep.n_saved(),
                                   (done + nsp - 1) // nsp if nsp > 0 else done)
                    # contract whole half rings as they fill; the rest at the end
                    if done >= a.req.prep.n_steps:
                        self._contract(a, produced)
                        self._finish(a)
                        retiring.append(a)
                    else:
                        full = (produced // self.ring) * self.ring
                        if full > a.done_rows:
                            self._contract(a, full)
                        still.append(a)
                active = still
            # bound the work queued ahead of the host
            while len(inflight) > 2:
                inflight.pop(0).synchronize()
            keep = []
            removed = False
            for a in retiring:
                if a.finished_event.query():
                    t_done = time.perf_counter() - t_start
                    lat.append(t_done - a.req.arrival)
                    voxel_steps += a.req.voxel_steps
                    records.append(dict(rid=a.req.rid, n_voxel=a.req.prep.n_voxel,
                                        steps=a.req.prep.n_steps, model=a.req.model,
                                        electrodes=len(a.req.electrodes),
                                        arrival=a.req.arrival, admitted=a.admitted,
                                        done=t_done, latency=t_done - a.req.arrival,
                                        **({"egm": a.results["egm"], "lat": a.results["lat"]}
                                           if self.keep else {})))
                    self.arena.remove(a.slot)
                    removed = True
                    a.lf.close()
                    self.in_use -= a.req.footprint_bytes(2 * self.ring)
                    if progress:
                        progress(records[-1])
                else:
                    keep.append(a)
            retiring = keep
            if removed:
                # close the ranges just freed, so that the launches cover the resident
                # bricks rather than the arena's high-water mark
                self.brick_moves += self.arena.compact()
            if not active and retiring:
                time.sleep(0.0005)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t_start
        la = np.asarray(lat)
        return ServeStats(lat, voxel_steps, wall, len(records), len(records),
                          float(np.percentile(la, 50)) if len(la) else float("nan"),
                          float(np.percentile(la, 99)) if len(la) else float("nan"),
                          voxel_steps / wall, records, self.brick_moves, self.voxel_moves)

    def _finish(self, a: _Active):
        torch = self.torch
        runtime._lib.check(self.arena.lib.phs_egm_normalize(
            runtime._lib.ptr(a.egm), a.egm.shape[0], a.egm.shape[1], 1.0, runtime._lib.stream_ptr()),
            "phs_egm_normalize")
        o = self.arena.slots[a.slot]
        perm = torch.from_numpy(o.perm).to("cuda")
        lat = self.arena._scatter(o.lat, perm, self.arena.lib.phs_scatter_rows_i32)
        if self.keep:
            a.results["egm"] = a.egm.to("cpu", non_blocking=True)
            a.results["lat"] = lat.to("cpu", non_blocking=True)
        ev = torch.cuda.Event()
        ev.record()
        a.finished_event = ev
