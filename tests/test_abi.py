"""CPU checks of the C ABI: the library loads, exports every symbol include/phs.h
declares, validates arguments before touching the device, and the host-only
brick planner is correct."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

from pyheartsim_amd import _lib, geometry as geo, runtime

HEADER = Path(__file__).resolve().parents[1] / "include" / "phs.h"


def declared():
    text = HEADER.read_text()
    return sorted(set(re.findall(r"^\w[\w\s\*]*?\b(phs_\w+)\s*\(", text, flags=re.M)))


def test_every_declared_symbol_is_exported_and_bound():
    lib = _lib.load()
    names = declared()
    assert len(names) >= 25
    for n in names:
        assert hasattr(lib, n), n
        assert n in _lib.SIGNATURES, f"{n} has no ctypes prototype"
    assert set(_lib.SIGNATURES) == set(names)
    assert lib.phs_version() >= 10000


def test_argument_errors_need_no_device():
    lib = _lib.load()
    h = C.c_void_p()
    assert lib.phs_plain_create(None, None, 1e-12, C.byref(h)) == -1
    assert b"desc" in lib.phs_last_error()
    cfg = _lib.ArenaConfig()
    cfg.max_voxels, cfg.max_bricks, cfg.max_sims = 100, 10, 1
    cfg.brick[0], cfg.brick[1], cfg.brick[2] = 4, 4, 3
    cfg.sweeps = 8
    assert lib.phs_arena_create(C.byref(cfg), C.byref(h)) == -1
    assert b"brick shape" in lib.phs_last_error()
    cfg.brick[2] = 4
    cfg.sweeps = 1
    assert lib.phs_arena_create(C.byref(cfg), C.byref(h)) == -1
    assert lib.phs_egm_run(None, None, 1, 1, None, 0, None) == -1
    moved = C.c_int32(7)
    assert lib.phs_arena_compact(None, 0.0, 1, C.byref(moved), None) == -1
    assert b"NULL" in lib.phs_last_error()


@pytest.mark.parametrize("brick", [(4, 4, 4), (8, 8, 1), (8, 4, 2), (4, 8, 2), (16, 4, 1)])
def test_brick_plan(brick):
    g = geo.shell((5.0, 4.0, 3.5), 1.2, 0.4, n_holes=2, hole_radius=1.5, seed=9)
    ijk = g["ijk"]
    perm, masks = runtime.brick_plan(ijk, g["neighbor_id_2d"], brick)
    n = ijk.shape[0]
    assert np.array_equal(np.sort(perm), np.arange(n))
    counts = np.array([bin(int(m)).count("1") for m in masks])
    assert counts.sum() == n and counts.min() >= 1
    # bricks are contiguous and in Morton order; within a brick voxels follow lane order
    bx, by, bz = brick
    rel = ijk[perm] - ijk.min(axis=0)
    bc = rel // np.array(brick)
    lane = rel[:, 0] % bx + bx * (rel[:, 1] % by + by * (rel[:, 2] % bz))
    starts = np.concatenate([[0], np.cumsum(counts)[:-1]])
    for b in range(len(masks)):
        s, e = starts[b], starts[b] + counts[b]
        assert (bc[s:e] == bc[s]).all()
        assert np.all(np.diff(lane[s:e]) > 0)
        assert int(masks[b]) == sum(1 << int(l) for l in lane[s:e])


def test_brick_plan_rejects_inconsistent_neighbours():
    g = geo.shell((4.0, 3.0, 3.0), 1.0, 0.5, n_holes=0, seed=1)
    nb = g["neighbor_id_2d"].copy()
    i = np.nonzero(nb[:, 0] >= 0)[0][0]
    nb[i, 0] = -1                      # a cut link is part of the table's operator
    runtime.brick_plan(g["ijk"], nb)
    nb[i, 0] = nb[i, 1] if nb[i, 1] >= 0 else (i + 1) % len(nb)   # the wrong voxel
    with pytest.raises(_lib.PhsError, match="neighbour table entry") as e:
        runtime.brick_plan(g["ijk"], nb)
    assert e.value.code == _lib.PHS_ERR_UNSUPPORTED
    with pytest.raises(_lib.PhsError, match="duplicate") as e:
        runtime.brick_plan(np.concatenate([g["ijk"], g["ijk"][:1]]))
    assert e.value.code == _lib.PHS_ERR_ARG
