"""ctypes wrapper of oracle/libcpu_oracle.so (cpu_oracle.c) -- TEST INFRASTRUCTURE ONLY.

The C/OpenMP FP64 restatement of legacy/compute_simulation_cpu.py with the
reference's BiCGSTAB stop rule.  Used as the timed CPU baseline on the host
cores and checked against the NumPy oracle in tests/test_cpu_oracle.py.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

from . import sim_oracle

HERE = Path(__file__).resolve().parent
LIB = HERE / "libcpu_oracle.so"


class _Desc(C.Structure):
    _fields_ = [
        ("n", C.c_int32), ("model", C.c_int32), ("p", C.c_double * 5),
        ("c", C.c_void_p), ("tb", C.c_void_p), ("nb", C.c_void_p),
        ("s1", C.c_void_p), ("s2", C.c_void_p),
        ("dt", C.c_double), ("delta", C.c_double),
        ("rtol", C.c_double), ("atol", C.c_double), ("maxiter", C.c_int32),
    ]


def build():
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


_lib = None


def load():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        _lib = C.CDLL(str(LIB))
        _lib.cpu_simulate.restype = C.c_int
        _lib.cpu_simulate.argtypes = [C.POINTER(_Desc), C.c_void_p, C.c_int64, C.c_void_p,
                                      C.c_void_p, C.c_int, C.c_int]
    return _lib


class CpuSim:
    """State and inputs of one simulation for the C restatement."""

    def __init__(self, P_2d, geometry_data, simulation_parameters, arrhythmia_parameters,
                 rtol=1e-5, atol=1e-6, maxiter=1000):
        sim, arr = simulation_parameters, arrhythmia_parameters
        P = np.asarray(P_2d, dtype=np.float64)
        n = P.shape[0]
        nf = np.asarray(arr["node_flag"])
        self.n = n
        self.c = np.ascontiguousarray(np.where(nf == 4, 0.0, P[:, 20]))
        self.tb = np.ascontiguousarray((nf == 3).astype(np.uint8))
        self.nb = np.ascontiguousarray(np.asarray(geometry_data["neighbor_id_2d"])[:, :6], np.int32)
        s1 = np.zeros(n, np.uint8)
        s1[np.atleast_1d(np.asarray(arr["s1_pacing_voxel_id"], dtype=np.int64))] = 1
        s2 = np.zeros(n, np.uint8)
        s2ids = np.atleast_1d(np.asarray(arr.get("s2_pacing_voxel_id", []), dtype=np.int64))
        s2[s2ids] = 1
        self.s1, self.s2 = s1, s2
        d = _Desc()
        d.n = n
        d.model = int(sim["heart_model_flag"])
        for i in range(5):
            d.p[i] = float(P[0, 15 + i])
        d.c, d.tb, d.nb = self.c.ctypes.data, self.tb.ctypes.data, self.nb.ctypes.data
        d.s1, d.s2 = self.s1.ctypes.data, self.s2.ctypes.data
        d.dt, d.delta = float(sim["dt"]), float(geometry_data["Delta"])
        d.rtol, d.atol, d.maxiter = rtol, atol, maxiter
        self.desc = d
        self.sim, self.arr = sim, arr
        self.u = np.zeros(n)
        self.h = np.ones(n) if d.model == 0 else np.zeros(n)
        self.step = 0

    def schedule(self, n_steps):
        s1_t, dur, s2_t = sim_oracle.pacing_constants(self.arr, self.sim)
        dt = self.sim["dt"]
        out = np.zeros(n_steps, np.uint8)
        for i in range(n_steps):
            t = (self.step + i) * dt
            on1, on2 = sim_oracle.stim_on(self.arr, self.sim, self.sim["arrhythmia_flag"], t, s1_t,
                                          dur, s2_t)
            out[i] = (1 if on1 else 0) | (2 if on2 else 0) | (4 if 0 <= t < 350 * dt else 0)
        return out

    def run(self, n_steps, threads=0, rebuild=True):
        """Advance n_steps steps on `threads` OpenMP threads (0 = the runtime's default).
        rebuild: form A = I - dt/2 L anew every step, as legacy/compute_simulation_cpu.py:131-133
        does; False forms it once per call.  Returns the count of solves that hit maxiter."""
        sched = self.schedule(n_steps)
        fails = load().cpu_simulate(C.byref(self.desc), sched.ctypes.data, n_steps,
                                    self.u.ctypes.data, self.h.ctypes.data, threads,
                                    1 if rebuild else 0)
        self.step += n_steps
        return fails


def threads_available() -> int:
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1
