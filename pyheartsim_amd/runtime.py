"""Host runtime around libphs: input preparation, the plain FP64 form, the
bricked FP32 arena.  Torch tensors are used only as device buffers.

``prepare`` turns the reference's five ``simulation_gpu.compute`` arguments
(simulation/simulation_gpu.py:261-357) into the flat arrays of
``phs_sim_desc``: uniform model parameters from P_2d[:, 15:20], the
conductivity (uniform, or per voxel), the scar / temporary-block / pacing-site
flag bytes, the FP64-evaluated per-step pacing schedule, and the sampling
constants of :286-290 and :355-356.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from .geometry import lattice_from_neighbors
from .simulation import pacing

LAT_THRESHOLD = 0.13


@dataclass
class Prepared:
    n_voxel: int
    model: int
    params: np.ndarray        # 5 uniform values of P_2d[:, 15:20]
    c: float
    c_voxel: np.ndarray | None
    dt: float
    delta: float
    time_scale: float
    n_steps: int
    steps_per_sample: int
    n_samples: int
    ijk: np.ndarray           # int32 N x 3
    nb: np.ndarray            # int32 N x 6
    vflags: np.ndarray        # uint8 N
    schedule: np.ndarray      # uint8 n_steps
    lat_threshold: float = LAT_THRESHOLD
    u0: np.ndarray | None = None
    h0: np.ndarray | None = None
    params_voxel: np.ndarray | None = None  # float64 5 x N when P_2d[:, 15:20] varies
    keep: list = field(default_factory=list)

    def physical_time(self) -> np.ndarray:
        """Label of sample i: model time of the step that produced it (:399)."""
        t = []
        save = 0
        for step in range(self.n_steps):
            ok = (step % self.steps_per_sample == 0) if self.steps_per_sample > 0 else True
            if ok and save < self.n_samples:
                t.append(step * self.dt * self.time_scale)
                save += 1
        return np.asarray(t, dtype=np.float32)

    def n_saved(self) -> int:
        if self.steps_per_sample > 0:
            return min(self.n_samples, (self.n_steps + self.steps_per_sample - 1) // self.steps_per_sample)
        return min(self.n_samples, self.n_steps)

    def desc(self) -> _lib.SimDesc:
        d = _lib.SimDesc()
        d.n_voxel = self.n_voxel
        d.model = self.model
        for i in range(5):
            d.params[i] = float(self.params[i])
        d.c = float(self.c)
        d.c_voxel = _lib.ptr(self.c_voxel)
        d.dt = float(self.dt)
        d.delta = float(self.delta)
        d.n_steps = self.n_steps
        d.steps_per_sample = self.steps_per_sample
        d.n_samples = self.n_samples
        d.ijk = _lib.ptr(self.ijk)
        d.nb = _lib.ptr(self.nb)
        d.nb_stride = 6
        d.vflags = _lib.ptr(self.vflags)
        d.schedule = _lib.ptr(self.schedule)
        d.u0 = _lib.ptr(self.u0)
        d.h0 = _lib.ptr(self.h0)
        d.lat_threshold = float(self.lat_threshold)
        d.params_voxel = _lib.ptr(self.params_voxel)
        return d


def lattice_coordinates(geometry_data) -> np.ndarray:
    if "ijk" in geometry_data:
        return np.ascontiguousarray(geometry_data["ijk"], dtype=np.int32)
    if "voxel" in geometry_data and len(geometry_data["voxel"]):
        v = np.asarray(geometry_data["voxel"], dtype=np.float64)
        d = float(geometry_data["Delta"])
        rel = (v - v.min(axis=0)) / d
        ijk = np.rint(rel)
        if np.max(np.abs(rel - ijk)) < 0.1:
            return ijk.astype(np.int32)
    return lattice_from_neighbors(geometry_data["neighbor_id_2d"]).astype(np.int32)


def prepare(n_voxel, P_2d, geometry_data, simulation_parameters, arrhythmia_parameters,
            need_lattice: bool = True) -> Prepared:
    P = np.asarray(P_2d, dtype=np.float64)
    if P.shape != (n_voxel, 21):
        raise ValueError(f"P_2d must be ({n_voxel}, 21), got {P.shape}")
    nb = np.asarray(geometry_data["neighbor_id_2d"])
    delta_k = (nb[:, :6] >= 0).astype(np.float64)
    if np.any(P[:, 6:15] != 0) or np.any(P[:, 0:6] != 4.0 * delta_k):
        # the C ABI carries no diffusion tensor: the device path is the reference's D = I
        # (fiber_flag 0, simulation/equation_parts.py:19, :32)
        raise _lib.PhsError(f"PHS_ERR_UNSUPPORTED ({_lib.PHS_ERR_UNSUPPORTED}): anisotropic "
                            "diffusion (P_2d columns 0-14 other than D = I) is not supported by "
                            "the device path", _lib.PHS_ERR_UNSUPPORTED)
    params = P[0, 15:20].copy()
    params_voxel = None
    if np.any(P[:, 15:20] != params[None, :]):
        # per-voxel cell parameters, read per voxel as simulation_gpu.py:29-41 does
        params_voxel = np.ascontiguousarray(P[:, 15:20].T, dtype=np.float64)
    sim, arr = simulation_parameters, arrhythmia_parameters
    node_flag = np.asarray(arr["node_flag"]) if "node_flag" in arr else np.zeros(n_voxel, np.int64)
    model = int(sim["heart_model_flag"])
    ts = sim["time_scale"]
    t_final = float(sim["t_final"])
    dt = sim["dt"]
    n_steps = pacing.total_steps(sim)
    nspm = int(1 / (dt * ts))
    n_samples = int(np.round(t_final)) if model == 0 else int(np.round(t_final * ts))

    vflags = np.zeros(n_voxel, dtype=np.uint8)
    c_col = P[:, 20]
    scar = (node_flag == 4) | (c_col == 0)
    vflags[scar] |= _lib.PHS_VF_SCAR
    vflags[node_flag == 3] |= _lib.PHS_VF_TEMP_BLOCK
    s1, s2 = pacing.site_ids(arr, sim, n_voxel)
    vflags[s1] |= _lib.PHS_VF_S1
    vflags[s2] |= _lib.PHS_VF_S2
    conducting = c_col[~scar]
    if conducting.size == 0:
        c, c_voxel = 0.0, None
    elif np.all(conducting == conducting[0]):
        c, c_voxel = float(conducting[0]), None
    else:
        c, c_voxel = float(conducting[0]), np.ascontiguousarray(c_col, dtype=np.float64)
    schedule = pacing.pacing_schedule(arr, sim, n_steps)
    ijk = lattice_coordinates(geometry_data) if need_lattice else np.zeros((n_voxel, 3), np.int32)
    return Prepared(
        n_voxel=n_voxel, model=model, params=params, c=c, c_voxel=c_voxel, dt=dt,
        delta=float(geometry_data["Delta"]), time_scale=ts, n_steps=n_steps,
        steps_per_sample=nspm, n_samples=n_samples, ijk=np.ascontiguousarray(ijk, np.int32),
        nb=np.ascontiguousarray(nb[:, :6], dtype=np.int32), vflags=vflags,
        schedule=np.ascontiguousarray(schedule, np.uint8), params_voxel=params_voxel)


def brick_plan(ijk, nb=None, brick=(4, 4, 4)):
    """Host-only Morton brick plan (phs_plan_build): (perm, masks[n_bricks])."""
    lib = _lib.load()
    ijk = np.ascontiguousarray(ijk, dtype=np.int32)
    n = ijk.shape[0]
    nb6 = None if nb is None else np.ascontiguousarray(np.asarray(nb)[:, :6], dtype=np.int32)
    b = np.ascontiguousarray(brick, dtype=np.int32)
    perm = np.empty(n, dtype=np.int32)
    masks = np.empty(n, dtype=np.uint64)
    nbk = C.c_int32()
    _lib.check(lib.phs_plan_build(n, _lib.ptr(ijk), _lib.ptr(nb6), 6, _lib.ptr(b), _lib.ptr(perm),
                                  C.byref(nbk), _lib.ptr(masks)), "phs_plan_build")
    return perm, masks[:nbk.value]


# --------------------------------------------------------------------- plain FP64 form
class PlainSim:
    """Plain FP64 form on one GPU (phs_plain_*).  Outputs in caller voxel order."""

    def __init__(self, prep: Prepared, tol: float = 1e-15, samples: bool = True,
                 sample_h: bool = True, sample_fp64: bool = True, max_beats: int = 7,
                 probes=None, device=None):
        import torch
        self.torch = torch
        self.prep = prep
        self.lib = _lib.load()
        dev = device or torch.device("cuda")
        N = prep.n_voxel
        T = prep.n_saved()
        dtype = torch.float64 if sample_fp64 else torch.float32
        self.su = torch.empty((T, N), dtype=dtype, device=dev) if samples else None
        self.sh = torch.empty((T, N), dtype=dtype, device=dev) if (samples and sample_h) else None
        self.lat = torch.empty((max_beats, N), dtype=torch.int32, device=dev)
        self.probes = None if probes is None else np.ascontiguousarray(probes, dtype=np.int32)
        self.probe_u = (torch.empty((prep.n_steps, len(self.probes)), dtype=dtype, device=dev)
                        if self.probes is not None else None)
        out = _lib.Outputs()
        out.samples_u = _lib.ptr(self.su)
        out.samples_h = _lib.ptr(self.sh)
        out.sample_fp64 = 1 if sample_fp64 else 0
        out.lat = _lib.ptr(self.lat)
        out.max_beats = max_beats
        out.probe_u = _lib.ptr(self.probe_u)
        out.probe_ids = _lib.ptr(self.probes)
        out.n_probes = 0 if self.probes is None else len(self.probes)
        self.handle = C.c_void_p()
        desc = prep.desc()
        _lib.check(self.lib.phs_plain_create(C.byref(desc), C.byref(out), tol, C.byref(self.handle)),
                   "phs_plain_create")
        self.steps_done = 0

    def info(self):
        sw, lo, hi = C.c_int32(), C.c_double(), C.c_double()
        _lib.check(self.lib.phs_plain_info(self.handle, C.byref(sw), C.byref(lo), C.byref(hi)))
        return sw.value, lo.value, hi.value

    def run(self, n_steps: int | None = None, stream=None):
        n = self.prep.n_steps - self.steps_done if n_steps is None else n_steps
        _lib.check(self.lib.phs_plain_run(self.handle, n, _lib.stream_ptr(stream)), "phs_plain_run")
        self.steps_done = min(self.prep.n_steps, self.steps_done + n)

    def nonfinite_step(self) -> int:
        """First step whose new u was NaN or inf at any voxel, or -1 (phs_plain_status)."""
        v = C.c_int64()
        _lib.check(self.lib.phs_plain_status(self.handle, C.byref(v), _lib.stream_ptr()),
                   "phs_plain_status")
        return v.value

    def state(self):
        torch = self.torch
        u = torch.empty(self.prep.n_voxel, dtype=torch.float64, device="cuda")
        h = torch.empty_like(u)
        _lib.check(self.lib.phs_plain_state(self.handle, _lib.ptr(u), _lib.ptr(h), _lib.stream_ptr()))
        return u, h

    def close(self):
        if self.handle:
            self.lib.phs_plain_destroy(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# --------------------------------------------------------------------- bricked FP32 arena
def stable_sweeps(lambda_max: float, guess: int = 2, refine: int = 0, refine_sweeps: int = 8):
    """(fewest Chebyshev sweeps whose inexact CN step grows no mode, the arena's automatic
    count) for spectral bound lambda_max (phs_stable_sweeps)."""
    lib = _lib.load()
    s, a = C.c_int32(), C.c_int32()
    _lib.check(lib.phs_stable_sweeps(float(lambda_max), guess, refine, refine_sweeps, C.byref(s),
                                     C.byref(a)), "phs_stable_sweeps")
    return s.value, a.value


def auto_sweeps(lambda_max: float, prep: "Prepared", guess: int = 2, refine: int = 0,
                refine_sweeps: int = 8):
    """(the arena's automatic sweep count for simulation `prep`, its time resolution s = dt
    times the cell model's steepest upstroke rate, the lower end of the Chebyshev interval
    [lo, lambda_max] of its first solve) (phs_auto_sweeps)."""
    lib = _lib.load()
    p = np.ascontiguousarray(prep.params, dtype=np.float64)
    m, s, lo = C.c_int32(), C.c_double(), C.c_double()
    _lib.check(lib.phs_auto_sweeps(float(lambda_max), guess, refine, refine_sweeps, prep.model,
                                   _lib.ptr(p), float(prep.dt), C.byref(m), C.byref(s),
                                   C.byref(lo)), "phs_auto_sweeps")
    return m.value, s.value, lo.value


def step_amplification(lambda_max: float, sweeps: int, guess: int = 2, refine: int = 0,
                       refine_sweeps: int = 8) -> float:
    """Largest per-step growth factor of any mode of the inexact CN step (phs_step_amplification)."""
    lib = _lib.load()
    r = C.c_double()
    _lib.check(lib.phs_step_amplification(float(lambda_max), guess, sweeps, refine, refine_sweeps,
                                          C.byref(r)), "phs_step_amplification")
    return r.value


@dataclass
class SolverConfig:
    # 0 = automatic: per simulation the fewest sweeps that keep the step stable for its
    # spectral bound with margin (solve error at the stiffest mode <= 1/4 of CN's damping,
    # stiff modes shrinking by e within 40 steps); the arena runs the largest over resident
    # simulations.  guess: warm start 0 none, 1 previous correction, 2 linear extrapolation
    sweeps: int = 0
    refine: int = 0
    refine_sweeps: int = 8
    brick: tuple = (4, 4, 4)
    xcd_remap: bool = True
    diag: bool = False
    bricks_per_wave: int = 1
    guess: int = 2


@dataclass
class SlotOutputs:
    su: object
    sh: object
    lat: object
    n_voxel: int
    perm: np.ndarray


class Arena:
    """Many simulations packed into one set of launches (phs_arena_*)."""

    def __init__(self, max_voxels: int, max_sims: int, solver: SolverConfig = SolverConfig(),
                 max_bricks: int | None = None):
        import torch
        self.torch = torch
        self.lib = _lib.load()
        self.solver = solver
        cfg = _lib.ArenaConfig()
        cfg.max_voxels = int(max_voxels)
        cfg.max_bricks = int(max_bricks if max_bricks is not None else max(64, max_voxels // 8))
        cfg.max_sims = int(max_sims)
        for i in range(3):
            cfg.brick[i] = int(solver.brick[i])
        cfg.sweeps = solver.sweeps
        cfg.refine = solver.refine
        cfg.refine_sweeps = solver.refine_sweeps
        cfg.xcd_remap = 1 if solver.xcd_remap else 0
        cfg.diag = 1 if solver.diag else 0
        cfg.bricks_per_wave = int(solver.bricks_per_wave)
        cfg.guess = int(solver.guess)
        self.cfg = cfg
        self.handle = C.c_void_p()
        _lib.check(self.lib.phs_arena_create(C.byref(cfg), C.byref(self.handle)), "phs_arena_create")
        self.slots: dict[int, SlotOutputs] = {}

    def add(self, prep: Prepared, samples: bool = True, sample_h: bool = False,
            max_beats: int = 7, ring: int = 0, stream=None) -> int:
        """Add a simulation.  ring > 0 keeps only `ring` sample rows, reused cyclically."""
        torch = self.torch
        N = prep.n_voxel
        T = ring if ring > 0 else prep.n_saved()
        su = torch.empty((T, N), dtype=torch.float32, device="cuda") if samples else None
        sh = torch.empty((T, N), dtype=torch.float32, device="cuda") if (samples and sample_h) else None
        lat = torch.empty((max_beats, N), dtype=torch.int32, device="cuda")
        out = _lib.Outputs()
        out.samples_u = _lib.ptr(su)
        out.samples_h = _lib.ptr(sh)
        out.lat = _lib.ptr(lat)
        out.max_beats = max_beats
        out.sample_ring = ring
        desc = prep.desc()
        slot = C.c_int32(-1)
        _lib.check(self.lib.phs_arena_add(self.handle, C.byref(desc), C.byref(out), C.byref(slot),
                                          _lib.stream_ptr(stream)), "phs_arena_add")
        perm = np.empty(N, dtype=np.int32)
        _lib.check(self.lib.phs_arena_perm(self.handle, slot.value, perm.ctypes.data))
        self.slots[slot.value] = SlotOutputs(su, sh, lat, N, perm)
        return slot.value

    def remove(self, slot: int, stream=None):
        _lib.check(self.lib.phs_arena_remove(self.handle, slot, _lib.stream_ptr(stream)),
                   "phs_arena_remove")
        self.slots.pop(slot, None)

    def compact(self, max_waste: float = 0.05, voxels: bool = False, stream=None) -> int:
        """When the launches cover more than (1 + max_waste) times the resident bricks,
        slide the simulations' brick records down until they are contiguous; with `voxels`,
        first slide their voxel data down too, so that the free voxels form one range
        (phs_arena_compact).  Returns the number of range moves.  Results are unchanged,
        bit for bit."""
        moved = C.c_int32()
        _lib.check(self.lib.phs_arena_compact(self.handle, float(max_waste), 1 if voxels else 0,
                                              C.byref(moved), _lib.stream_ptr(stream)),
                   "phs_arena_compact")
        return moved.value

    def free_capacity(self):
        """(free voxels, free bricks, free slots) of the arena, summed over its free ranges."""
        c = self.counts()
        return (self.cfg.max_voxels - c["voxels"], self.cfg.max_bricks - c["bricks"],
                self.cfg.max_sims - len(self.slots))

    def step(self, n: int, graph: bool = False, stream=None):
        fn = self.lib.phs_arena_step_graph if graph else self.lib.phs_arena_step
        _lib.check(fn(self.handle, n, _lib.stream_ptr(stream)), "phs_arena_step")

    def prepare_graph(self, n: int):
        """Capture the n-step graphs of both launch parities now, so that replays of n steps
        do no capture (phs_arena_graph_prepare)."""
        _lib.check(self.lib.phs_arena_graph_prepare(self.handle, n), "phs_arena_graph_prepare")

    STEP_KINDS = ("rhs", "first_sweep", "mid_sweep", "last_sweep", "residual", "refine_sweep")

    def step_marked(self, stream=None):
        """One eager step with HIP events around each launch (phs_arena_step_marked)."""
        _lib.check(self.lib.phs_arena_step_marked(self.handle, _lib.stream_ptr(stream)),
                   "phs_arena_step_marked")

    def step_timing(self):
        """Launch durations (ms, HIP events) of the latest marked step, in launch order:
        [(kind, ms), ...] (phs_arena_step_timing)."""
        cap = 512
        ms = (C.c_double * cap)()
        kinds = (C.c_int32 * cap)()
        n = C.c_int32()
        _lib.check(self.lib.phs_arena_step_timing(self.handle, cap, ms, kinds, C.byref(n)),
                   "phs_arena_step_timing")
        return [(self.STEP_KINDS[kinds[i]], ms[i]) for i in range(min(n.value, cap))]

    def steps_done(self, slot: int) -> int:
        s = C.c_int64()
        _lib.check(self.lib.phs_arena_slot_info(self.handle, slot, C.byref(s), None, None, None))
        return s.value

    def slot_info(self, slot: int):
        s, n, nb, off = C.c_int64(), C.c_int32(), C.c_int32(), C.c_int64()
        _lib.check(self.lib.phs_arena_slot_info(self.handle, slot, C.byref(s), C.byref(n),
                                                C.byref(nb), C.byref(off)))
        return dict(steps_done=s.value, n_voxel=n.value, n_bricks=nb.value, voxel_offset=off.value)

    def bounds(self, slot: int):
        lo, hi = C.c_double(), C.c_double()
        _lib.check(self.lib.phs_arena_slot_bounds(self.handle, slot, C.byref(lo), C.byref(hi)))
        return lo.value, hi.value

    def nonfinite_step(self, slot: int) -> int:
        """First step whose new u was NaN or inf at any voxel of the slot, or -1
        (phs_arena_slot_status)."""
        v = C.c_int64()
        _lib.check(self.lib.phs_arena_slot_status(self.handle, slot, C.byref(v), _lib.stream_ptr()),
                   "phs_arena_slot_status")
        return v.value

    def last_residual(self, slot: int) -> float:
        r = C.c_double()
        _lib.check(self.lib.phs_arena_last_residual(self.handle, slot, C.byref(r), _lib.stream_ptr()),
                   "phs_arena_last_residual")
        return r.value

    def sweeps(self) -> int:
        """Sweeps of the first CN solve that the next step runs."""
        s = C.c_int32()
        _lib.check(self.lib.phs_arena_sweeps(self.handle, C.byref(s)), "phs_arena_sweeps")
        return s.value

    def profile(self, reps: int = 50, stream=None):
        """Average kernel durations (ms): reaction, rhs, sweep (HIP events)."""
        ms = (C.c_double * 3)()
        _lib.check(self.lib.phs_arena_profile(self.handle, reps, ms, _lib.stream_ptr(stream)),
                   "phs_arena_profile")
        return dict(reaction=ms[0], rhs=ms[1], sweep=ms[2])

    def counts(self):
        v, b, f, lb = C.c_int64(), C.c_int64(), C.c_int64(), C.c_int64()
        _lib.check(self.lib.phs_arena_counts(self.handle, C.byref(v), C.byref(b), C.byref(f),
                                             C.byref(lb)))
        return dict(voxels=v.value, bricks=b.value, vf_voxels=f.value, launch_bricks=lb.value)

    def state(self, slot: int):
        """(u, h) of a slot in caller order.  Between steps u is the next step's u* (its
        reaction runs at the end of each step); after the last step, the final state."""
        torch = self.torch
        n = self.slots[slot].n_voxel
        u = torch.empty(n, dtype=torch.float32, device="cuda")
        h = torch.empty(n, dtype=torch.float32, device="cuda")
        _lib.check(self.lib.phs_arena_state(self.handle, slot, _lib.ptr(u), _lib.ptr(h),
                                            _lib.stream_ptr()), "phs_arena_state")
        return u, h

    def _scatter(self, src, perm_dev, fn):
        torch = self.torch
        dst = torch.empty_like(src)
        rows, n = src.shape
        _lib.check(fn(_lib.ptr(dst), _lib.ptr(src), _lib.ptr(perm_dev), rows, n, _lib.stream_ptr()),
                   "scatter")
        return dst

    def lat(self, slot: int):
        """Activation times (max_beats x N int32, -1 = none) in caller voxel order."""
        o = self.slots[slot]
        perm = self.torch.from_numpy(o.perm).to("cuda")
        return self._scatter(o.lat, perm, self.lib.phs_scatter_rows_i32)

    def outputs(self, slot: int):
        """(samples_u, samples_h, lat) in caller voxel order (device tensors)."""
        torch = self.torch
        o = self.slots[slot]
        perm = torch.from_numpy(o.perm).to("cuda")
        su = self._scatter(o.su, perm, self.lib.phs_scatter_rows_f32) if o.su is not None else None
        sh = self._scatter(o.sh, perm, self.lib.phs_scatter_rows_f32) if o.sh is not None else None
        lat = self._scatter(o.lat, perm, self.lib.phs_scatter_rows_i32)
        return su, sh, lat

    def close(self):
        if self.handle:
            self.lib.phs_arena_destroy(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
