"""Autotuner for the bricked FP32 arena: brick shape, bricks per wavefront and the
XCD-contiguous workgroup mapping.

Candidates are first ranked on the host by brick fill (occupied cells / 64 per
brick, from the host planner); the best few shapes are then timed on the device:
each candidate arena holds the same simulations and runs `steps` graph-replayed
steps after a warm-up, and the fastest configuration wins.  The choice does not
change results beyond FP32 summation order (each voxel's arithmetic is the same;
only the neighbour-sum order of in-brick vs cross-brick terms can differ with the
brick shape).
"""
from __future__ import annotations

import dataclasses
import time

import numpy as np

from . import runtime

SHAPES = ((4, 4, 4), (8, 4, 2), (4, 8, 2), (8, 8, 1), (16, 4, 1))


def brick_fill(prep: runtime.Prepared, shape) -> float:
    perm, masks = runtime.brick_plan(prep.ijk, None, shape)
    return len(perm) / (64.0 * len(masks))


def rank_shapes(preps, top: int = 3):
    scores = []
    for s in SHAPES:
        n = sum(p.n_voxel for p in preps)
        occ = sum(brick_fill(p, s) * p.n_voxel for p in preps) / n
        scores.append((occ, s))
    scores.sort(reverse=True)
    return [s for _, s in scores[:top]], {s: o for o, s in scores}


def tune(preps, base: runtime.SolverConfig, steps: int = 200, warmup: int = 50,
         graph_steps: int = 50, nbs=(1, 2), top_shapes: int = 2, xcd=(True,)):
    """Time each candidate on the device; returns (best SolverConfig, table of results)."""
    import torch
    shapes, fill = rank_shapes(preps, top_shapes)
    n_vox = sum(p.n_voxel for p in preps)
    table = []
    best = None
    for shape in shapes:
        for nb in nbs:
            for x in xcd:
                cfg = dataclasses.replace(base, brick=tuple(shape), bricks_per_wave=nb, xcd_remap=x)
                arena = runtime.Arena(n_vox, len(preps), cfg)
                for p in preps:
                    arena.add(p, samples=False)
                g = max(1, min(graph_steps, warmup))
                for _ in range(max(1, warmup // g)):
                    arena.step(g, graph=True)
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                n = 0
                while n < steps:
                    arena.step(g, graph=True)
                    n += g
                torch.cuda.synchronize()
                dt = time.perf_counter() - t0
                rate = n_vox * n / dt
                prof = arena.profile(20)
                table.append(dict(brick=list(shape), bricks_per_wave=nb, xcd_remap=x,
                                  fill=fill[shape], voxel_steps_per_s=rate, kernels_ms=prof))
                if best is None or rate > best[0]:
                    best = (rate, cfg)
                arena.close()
                del arena
    return best[1], table


def fill_table(prep: runtime.Prepared):
    return {s: brick_fill(prep, s) for s in SHAPES}


def lat_error_steps(lat, ref_lat):
    """Max |LAT - LAT_ref| in steps over voxels and beats; inf if activation differs."""
    a = np.asarray(lat, dtype=np.int64)
    b = np.asarray(ref_lat, dtype=np.int64)
    if not np.array_equal(a >= 0, b >= 0):
        return float("inf")
    m = a >= 0
    return float(np.max(np.abs(a[m] - b[m]))) if m.any() else 0.0


def select_sweeps(prep: runtime.Prepared, ref_lat, base: runtime.SolverConfig,
                  candidates=(4, 5, 6, 8, 10, 12, 16), max_lat_err: float = 1.0,
                  graph_steps: int = 50):
    """Smallest Chebyshev sweep count whose full-length run keeps every voxel's LAT within
    `max_lat_err` steps of `ref_lat` (the plain FP64 form's LAT, caller voxel order).
    Counts below the stability limit of the simulation's spectral bound are skipped (the
    arena refuses them).  Returns (sweeps or None, table)."""
    table = []
    chosen = None
    probe = runtime.Arena(prep.n_voxel, 1, dataclasses.replace(base, sweeps=0))
    lmax = probe.bounds(probe.add(prep, samples=False, max_beats=1))[1]
    probe.close()
    stable, _ = runtime.stable_sweeps(lmax, base.guess, base.refine, base.refine_sweeps)
    for m in sorted(candidates):
        if m < stable:
            table.append(dict(sweeps=m, guess=base.guess, lat_err_steps=None, unstable=True))
            continue
        cfg = dataclasses.replace(base, sweeps=m)
        arena = runtime.Arena(prep.n_voxel, 1, cfg)
        slot = arena.add(prep, samples=False)
        while arena.steps_done(slot) < prep.n_steps:
            arena.step(graph_steps, graph=True)
        _, _, lat = arena.outputs(slot)
        err = lat_error_steps(lat.cpu().numpy(), ref_lat)
        table.append(dict(sweeps=m, guess=cfg.guess, lat_err_steps=err))
        arena.close()
        if err <= max_lat_err and chosen is None:
            chosen = m
    return chosen, table


if __name__ == "__main__":
    # python -m pyheartsim_amd.autotune [n]: tune over n copies (default 16) of the 0.4 mm
    # shell with different pacing sites, every shape and both workgroup mappings
    import json
    import sys

    from . import configs
    from . import geometry as geo
    g = geo.shell((25.0, 20.0, 18.0), 3.0, 0.4, 4, 6.0, 0)
    preps = []
    for s in range(int(sys.argv[1]) if len(sys.argv) > 1 else 16):
        w = configs.shell_ms(seed=s, geometry=g)
        P = configs.build_P_2d(g, w["sim"], w["hm"])
        preps.append(runtime.prepare(P.shape[0], P, g, w["sim"], w["arr"]))
    best, table = tune(preps, runtime.SolverConfig(), top_shapes=len(SHAPES),
                       xcd=(True, False))
    for row in table:
        print(json.dumps(row), flush=True)
    print("best", best, flush=True)
