"""phs_egm_run_noalloc (split-K workspace in a static device array, no allocation) against
phs_egm_run, through the C ABI.

Bars: bitwise-equal electrograms for both precisions and electrode counts that use the
one-, two- and four-tile workgroups, on two calls in a row with the second input smaller; a
request beyond the static array fails with an error before any launch, and the entry point
keeps working after it.
"""
import ctypes as C

import numpy as np
import pytest

from tests.helpers import small_shell

pytestmark = pytest.mark.gpu


def _run(lib, fn, lf, samples, out, precision):
    from pyheartsim_amd import _lib
    T, N = samples.shape
    return fn(lf.handle, _lib.ptr(samples), T, N, _lib.ptr(out), precision, _lib.stream_ptr())


@pytest.mark.parametrize("E,precision", [(64, 0), (64, 1), (100, 0), (200, 0)])
def test_noalloc_matches_allocating_entry_point(cuda, E, precision):
    torch = cuda
    from pyheartsim_amd import _lib
    from pyheartsim_amd.simulation.unipolar_electrogram_gpu import LeadField
    lib = _lib.load()
    w = small_shell()
    g = w["geometry"]
    N = g["voxel"].shape[0]
    rng = np.random.default_rng(E + precision)
    el = rng.uniform(-5, 5, size=(E, 3))
    lf = LeadField(el, g["voxel"], None, w["hm"]["c_voxel"], g["Delta"], g["neighbor_id_2d"])
    for T in (300, 37):   # two calls in a row, the second smaller
        s = torch.from_numpy(rng.uniform(-0.1, 1.1, size=(T, N)).astype(np.float32)).cuda()
        ref = torch.empty((T, E), device="cuda")
        got = torch.full((T, E), float("nan"), device="cuda")
        _lib.check(_run(lib, lib.phs_egm_run, lf, s, ref, precision), "phs_egm_run")
        _lib.check(_run(lib, lib.phs_egm_run_noalloc, lf, s, got, precision),
                   "phs_egm_run_noalloc")
        assert np.array_equal(got.cpu().numpy(), ref.cpu().numpy()), T
    lf.close()


def test_noalloc_rejects_requests_beyond_the_scratch(cuda):
    torch = cuda
    from pyheartsim_amd import _lib
    from pyheartsim_amd.simulation.unipolar_electrogram_gpu import LeadField
    lib = _lib.load()
    cap = C.c_int64()
    _lib.check(lib.phs_egm_scratch_bytes(C.byref(cap)))
    assert cap.value == 64 << 20
    w = small_shell()
    g = w["geometry"]
    N = g["voxel"].shape[0]
    el = np.random.default_rng(1).uniform(-5, 5, size=(256, 3))
    lf = LeadField(el, g["voxel"], None, w["hm"]["c_voxel"], g["Delta"], g["neighbor_id_2d"])
    # 70000 samples x 256 electrodes x 4 B > 64 MiB of workspace even with a single K split
    T = 70000
    s = torch.zeros((T, N), device="cuda")
    out = torch.zeros((T, 256), device="cuda")
    rc = _run(lib, lib.phs_egm_run_noalloc, lf, s, out, 0)
    assert rc == -1
    assert "static scratch" in lib.phs_last_error().decode()
    # a valid call still works after the rejection, and equals the allocating entry point
    small = s[:64]
    ref = torch.empty((64, 256), device="cuda")
    got = torch.empty((64, 256), device="cuda")
    small.copy_(torch.rand((64, N), device="cuda"))
    _lib.check(_run(lib, lib.phs_egm_run, lf, small, ref, 0))
    _lib.check(_run(lib, lib.phs_egm_run_noalloc, lf, small, got, 0))
    assert np.array_equal(got.cpu().numpy(), ref.cpu().numpy())
    lf.close()
