"""Drop-in for simulation/unipolar_electrogram_gpu.py.

``compute`` keeps the reference signature and return value
(simulation/unipolar_electrogram_gpu.py:71, :152): a (T x E) float32,
magnitude-normalised electrogram.  ``batch_size`` is accepted for
compatibility; the device path contracts all samples in one launch.

Keyword-only extras:
    precision="bf16x3"  BF16 hi/lo split MFMA, FP32 accumulation (default)
              "fp32"    FP32 MFMA
              "bf16"    single BF16 product (fastest, least accurate)
    normalize=False     return the raw (un-normalised) electrogram
    device_out=True     return a torch device tensor
``action_potential`` may be a NumPy array or a torch CUDA tensor (T x N).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from .. import _lib
from .equation_parts import IsotropicFibers, tensor_components

PRECISION = {"bf16x3": 0, "fp32": 1, "bf16": 2}


class LeadField:
    """Device lead field Wt (E x N) of one geometry and electrode set (phs_egm_*).

    ``perm`` maps sample columns to caller voxel ids when the samples are in
    another voxel order (column i holds voxel perm[i]).
    """

    def __init__(self, electrode_xyz, voxel, D0, c_voxel, Delta, neighbor_id_2d, perm=None,
                 stream=None):
        self.lib = _lib.load()
        voxel = np.ascontiguousarray(voxel, dtype=np.float64)
        el = np.ascontiguousarray(electrode_xyz, dtype=np.float64)
        c = np.ascontiguousarray(c_voxel, dtype=np.float64)
        N = voxel.shape[0]
        self.N, self.E = N, el.shape[0]
        if D0 is None or isinstance(D0, IsotropicFibers):
            D = None
        else:
            D = np.ascontiguousarray(tensor_components(D0, N))
            if np.all(D == np.array([1, 0, 0, 0, 1, 0, 0, 0, 1.0])[:, None]):
                D = None
        nb = np.ascontiguousarray(np.asarray(neighbor_id_2d)[:, :6], dtype=np.int32)
        p = None if perm is None else np.ascontiguousarray(perm, dtype=np.int32)
        self.handle = C.c_void_p()
        _lib.check(self.lib.phs_egm_create(N, self.E, _lib.ptr(voxel), _lib.ptr(el), _lib.ptr(c),
                                           _lib.ptr(D), _lib.ptr(nb), 6, float(Delta), _lib.ptr(p),
                                           C.byref(self.handle), _lib.stream_ptr(stream)),
                   "phs_egm_create")

    def run(self, samples, precision: str = "bf16x3", normalize: bool = True, stream=None,
            out=None):
        """samples: CUDA float32 tensor (T x stride), first N columns used.  `out`: optional
        contiguous (T x E) float32 destination."""
        import torch
        T, stride = samples.shape
        if out is None:
            out = torch.empty((T, self.E), dtype=torch.float32, device=samples.device)
        sp = _lib.stream_ptr(stream)
        _lib.check(self.lib.phs_egm_run(self.handle, _lib.ptr(samples), T, stride, _lib.ptr(out),
                                        PRECISION[precision], sp), "phs_egm_run")
        if normalize:
            _lib.check(self.lib.phs_egm_normalize(_lib.ptr(out), T, self.E, 1.0, sp),
                       "phs_egm_normalize")
        return out

    def close(self):
        if self.handle:
            self.lib.phs_egm_destroy(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class StreamContraction:
    """Contraction of a few new sample rows (1..16 per job) of many simulations in one launch
    (phs_egm_stream): each job is (LeadField, samples (T x stride) CUDA view, out (T x E)
    contiguous CUDA view).  The workspace is a device buffer kept between calls; `reserve`
    sizes it ahead of time so that later calls allocate nothing."""

    def __init__(self):
        self.lib = _lib.load()
        self.ws = None

    def _table(self, jobs):
        arr = (_lib.EgmJob * len(jobs))()
        for d, (lf, samples, out) in zip(arr, jobs):
            if samples.dim() != 2 or samples.stride(1) != 1:
                raise ValueError("samples must be a (T x stride) row-major view")
            T, stride = samples.shape[0], samples.stride(0)
            if out.shape[0] != T or out.shape[1] != lf.E or not out.is_contiguous():
                raise ValueError("out must be a contiguous (T x E) view")
            d.lead_field = lf.handle
            d.samples = _lib.ptr(samples)
            d.n_time = T
            d.sample_stride = stride
            d.egm = _lib.ptr(out)
        return arr

    def _need(self, arr) -> int:
        b = C.c_int64()
        _lib.check(self.lib.phs_egm_stream_workspace(arr, len(arr), C.byref(b)),
                   "phs_egm_stream_workspace")
        return b.value

    def reserve(self, jobs):
        import torch
        need = self._need(self._table(jobs))
        if self.ws is None or self.ws.numel() * 4 < need:
            self.ws = torch.empty(max(need // 4, 1), dtype=torch.float32, device="cuda")

    def run(self, jobs, stream=None):
        if not jobs:
            return
        arr = self._table(jobs)
        need = self._need(arr)
        if self.ws is None or self.ws.numel() * 4 < need:
            self.reserve(jobs)
        _lib.check(self.lib.phs_egm_stream(arr, len(arr), _lib.ptr(self.ws), self.ws.numel() * 4,
                                           _lib.stream_ptr(stream)), "phs_egm_stream")


def compute(electrode_xyz, voxel, D0, c_voxel, action_potential, Delta, neighbor_id_2d,
            batch_size=50, *, precision: str = "bf16x3", normalize: bool = True,
            device_out: bool = False):
    import torch
    if isinstance(action_potential, torch.Tensor):
        ap = action_potential.to(device="cuda", dtype=torch.float32).contiguous()
    else:
        ap = torch.from_numpy(np.ascontiguousarray(action_potential, dtype=np.float32)).to("cuda")
    lf = LeadField(electrode_xyz, voxel, D0, c_voxel, Delta, neighbor_id_2d)
    out = lf.run(ap, precision=precision, normalize=normalize)
    lf.close()
    return out if device_out else out.cpu().numpy()
