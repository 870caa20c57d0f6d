"""ctypes binding of libphs.so (include/phs.h).

The library is built in-tree by ``python -m pyheartsim_amd.build``.  Loading
fails loudly when it is missing: there is no CPU fallback on the product path.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libphs.so"

PHS_ERR_ARG = -1
PHS_ERR_HIP = -2
PHS_ERR_OOM = -3
PHS_ERR_UNSUPPORTED = -4
PHS_ERR_CAPACITY = -5
PHS_ERR_STATE = -6

PHS_VF_SCAR = 0x01
PHS_VF_TEMP_BLOCK = 0x02
PHS_VF_S1 = 0x04
PHS_VF_S2 = 0x08


class SimDesc(C.Structure):
    _fields_ = [
        ("n_voxel", C.c_int32),
        ("model", C.c_int32),
        ("params", C.c_double * 5),
        ("c", C.c_double),
        ("c_voxel", C.c_void_p),
        ("dt", C.c_double),
        ("delta", C.c_double),
        ("n_steps", C.c_int64),
        ("steps_per_sample", C.c_int32),
        ("n_samples", C.c_int32),
        ("ijk", C.c_void_p),
        ("nb", C.c_void_p),
        ("nb_stride", C.c_int32),
        ("vflags", C.c_void_p),
        ("schedule", C.c_void_p),
        ("u0", C.c_void_p),
        ("h0", C.c_void_p),
        ("lat_threshold", C.c_double),
        ("params_voxel", C.c_void_p),
    ]


class Outputs(C.Structure):
    _fields_ = [
        ("samples_u", C.c_void_p),
        ("samples_h", C.c_void_p),
        ("sample_fp64", C.c_int32),
        ("lat", C.c_void_p),
        ("max_beats", C.c_int32),
        ("probe_u", C.c_void_p),
        ("probe_ids", C.c_void_p),
        ("n_probes", C.c_int32),
        ("sample_ring", C.c_int32),
    ]


class ArenaConfig(C.Structure):
    _fields_ = [
        ("max_voxels", C.c_int64),
        ("max_bricks", C.c_int64),
        ("max_sims", C.c_int32),
        ("brick", C.c_int32 * 3),
        ("sweeps", C.c_int32),
        ("refine", C.c_int32),
        ("refine_sweeps", C.c_int32),
        ("xcd_remap", C.c_int32),
        ("diag", C.c_int32),
        ("bricks_per_wave", C.c_int32),
        ("guess", C.c_int32),
    ]


class EgmJob(C.Structure):
    _fields_ = [
        ("lead_field", C.c_void_p),
        ("samples", C.c_void_p),
        ("n_time", C.c_int64),
        ("sample_stride", C.c_int64),
        ("egm", C.c_void_p),
    ]


_P = C.c_void_p
_I32 = C.c_int32
_I64 = C.c_int64
_D = C.c_double
_PI32 = C.POINTER(C.c_int32)
_PI64 = C.POINTER(C.c_int64)
_PD = C.POINTER(C.c_double)

# name -> (restype, argtypes); every symbol declared in include/phs.h
SIGNATURES = {
    "phs_version": (C.c_int, []),
    "phs_last_error": (C.c_char_p, []),
    "phs_device_count": (C.c_int, [_PI32]),
    "phs_stable_sweeps": (C.c_int, [_D, _I32, _I32, _I32, _PI32, _PI32]),
    "phs_step_amplification": (C.c_int, [_D, _I32, _I32, _I32, _I32, _PD]),
    "phs_auto_sweeps": (C.c_int, [_D, _I32, _I32, _I32, _I32, _P, _D, _PI32, _PD, _PD]),
    "phs_plan_build": (C.c_int, [_I32, _P, _P, _I32, _P, _P, _PI32, _P]),
    "phs_plain_create":(C.c_int, [C.POINTER(SimDesc), C.POINTER(Outputs), _D, C.POINTER(_P)]),
    "phs_plain_run": (C.c_int, [_P, _I64, _P]),
    "phs_plain_state": (C.c_int, [_P, _P, _P, _P]),
    "phs_plain_info": (C.c_int, [_P, _PI32, _PD, _PD]),
    "phs_plain_status": (C.c_int, [_P, _PI64, _P]),
    "phs_plain_destroy": (C.c_int, [_P]),
    "phs_arena_create": (C.c_int, [C.POINTER(ArenaConfig), C.POINTER(_P)]),
    "phs_arena_add": (C.c_int, [_P, C.POINTER(SimDesc), C.POINTER(Outputs), _PI32, _P]),
    "phs_arena_remove": (C.c_int, [_P, _I32, _P]),
    "phs_arena_compact": (C.c_int, [_P, _D, _I32, _PI32, _P]),
    "phs_arena_step": (C.c_int, [_P, _I32, _P]),
    "phs_arena_step_graph": (C.c_int, [_P, _I32, _P]),
    "phs_arena_graph_prepare": (C.c_int, [_P, _I32]),
    "phs_arena_step_marked": (C.c_int, [_P, _P]),
    "phs_arena_step_timing": (C.c_int, [_P, _I32, _PD, _PI32, _PI32]),
    "phs_arena_slot_info": (C.c_int, [_P, _I32, _PI64, _PI32, _PI32, _PI64]),
    "phs_arena_perm": (C.c_int, [_P, _I32, _P]),
    "phs_arena_state": (C.c_int, [_P, _I32, _P, _P, _P]),
    "phs_arena_slot_bounds": (C.c_int, [_P, _I32, _PD, _PD]),
    "phs_arena_slot_status": (C.c_int, [_P, _I32, _PI64, _P]),
    "phs_arena_last_residual": (C.c_int, [_P, _I32, _PD, _P]),
    "phs_arena_sweeps": (C.c_int, [_P, _PI32]),
    "phs_arena_profile": (C.c_int, [_P, _I32, _PD, _P]),
    "phs_arena_counts": (C.c_int, [_P, _PI64, _PI64, _PI64, _PI64]),
    "phs_arena_destroy": (C.c_int, [_P]),
    "phs_scatter_rows_f32": (C.c_int, [_P, _P, _P, _I64, _I64, _P]),
    "phs_scatter_rows_i32": (C.c_int, [_P, _P, _P, _I64, _I64, _P]),
    "phs_egm_create": (C.c_int, [_I32, _I32, _P, _P, _P, _P, _P, _I32, _D, _P, C.POINTER(_P), _P]),
    "phs_egm_run": (C.c_int, [_P, _P, _I64, _I64, _P, _I32, _P]),
    "phs_egm_destroy": (C.c_int, [_P]),
    "phs_egm_run_noalloc": (C.c_int, [_P, _P, _I64, _I64, _P, _I32, _P]),
    "phs_egm_scratch_bytes": (C.c_int, [_PI64]),
    "phs_egm_stream_workspace": (C.c_int, [_P, _I32, _PI64]),
    "phs_egm_stream": (C.c_int, [_P, _I32, _P, _I64, _P]),
    "phs_egm_normalize": (C.c_int, [_P, _I64, _I32, _D, _P]),
    "phs_egm_electrode_lat": (C.c_int, [_P, _I64, _I32, _P, _P]),
    "phs_conduction_velocity": (C.c_int, [_P, _P, _I32, _D, _P, _P, _P, _P]),
    "phs_ap_lat_workspace": (C.c_int, [_I64, _PI64]),
    "phs_ap_lat": (C.c_int, [_P, _I64, _I64, _I64, _D, _D, _D, _P, _P, _P, _P, _P]),
    "phs_ap_markers": (C.c_int, [_P, _I64, _I64, _I64, _D, _P, _P, _P, _P]),
}

_lib = None


class PhsError(RuntimeError):
    """A failed library call; `code` is its PHS_ERR_* return value (None when the failure
    was found before any call, e.g. a missing library)."""

    def __init__(self, message: str, code: int | None = None):
        super().__init__(message)
        self.code = code


def load():
    """Load libphs.so and declare every entry point; raise if it is absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise PhsError(f"{LIB_PATH} is missing: build it with `python -m pyheartsim_amd.build` "
                       "(the HIP library is required; there is no CPU fallback)")
    lib = C.CDLL(str(LIB_PATH))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc: int, what: str = ""):
    if rc != 0:
        msg = load().phs_last_error().decode(errors="replace")
        raise PhsError(f"{what} failed ({rc}): {msg}", rc)


def ptr(t) -> int:
    """Device or host address of a torch tensor / numpy array (0 for None)."""
    if t is None:
        return 0
    if hasattr(t, "data_ptr"):
        return t.data_ptr()
    return t.ctypes.data


def stream_ptr(stream=None) -> int:
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return s.cuda_stream
