"""CPU checks of bench.py's contract: more than one GPU is refused with a message, and the
HBM traffic figures come from the committed rocprofv3 summaries only for the workload they
were measured on."""
import subprocess
import sys
from pathlib import Path

import bench

ROOT = Path(__file__).resolve().parents[1]


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout.strip() == ""


def _traffic_lines():
    import json
    out = []
    for path in bench.PROFILE_TRAFFIC:
        try:
            with open(path) as f:
                out += [json.loads(l[len("traffic:"):]) for l in f if l.startswith("traffic:")]
        except OSError:
            pass
    return out


def test_profiled_traffic_matches_measured_arena_only():
    """A committed traffic summary is used only for an arena of the same voxels, bricks and
    brick-record size as the one it was measured on."""
    for rec in _traffic_lines():
        cnt = {"voxels": rec["voxels"], "bricks": rec["bricks"]}
        t = bench.profiled_traffic(cnt)
        if rec["record_bytes"] == bench.BRICK_RECORD_BYTES:
            assert t is not None and t > 16 * rec["voxels"]
        else:
            assert t is None or any(r["record_bytes"] == bench.BRICK_RECORD_BYTES and
                                    r["voxels"] == rec["voxels"] for r in _traffic_lines())
    assert bench.profiled_traffic({"voxels": 1000, "bricks": 20}) is None


def test_profiled_traffic_of_the_electrogram():
    for n, e, t in ((199446, 64, 2000), (1596158, 256, 2048)):
        rec = bench.egm_traffic(n, e, t)
        algorithmic = 4.0 * t * n + 4.0 * e * n + 4.0 * t * e
        assert rec is not None and algorithmic < rec < 1.3 * algorithmic
    assert bench.egm_traffic(199446, 256, 2000) is None


def test_sweep_bytes_follow_the_kernels():
    """Bytes per voxel of each Chebyshev sweep: the first reads r and writes d; k = 2 also reads
    r (d_1 = r / theta is not read); k >= 3 also d_k-1; the last reads u*, h, the flag byte
    and the kept corrections and writes u*, h and the corrections."""
    assert [bench.sweep_bytes(k, 6) for k in range(1, 6)] == [8, 12, 16, 16, 45]
    assert bench.sweep_bytes(4, 5, guess=1) == 37 and bench.sweep_bytes(4, 5, guess=0) == 29
    assert bench.sweep_bytes(2, 3) == 41          # last sweep at k = 2: no d_k-1
    total, per = bench.step_bytes({"voxels": 1, "bricks": 0}, 5)
    assert total == 12 + 8 + 12 + 16 + 45 and per["mid"] == 14.0


def test_algorithmic_bytes_match_the_measured_traffic():
    """The committed counter summary of the batched arena moves each kernel's algorithmic
    bytes within 5 % (the byte accounting is checked against the hardware, not restated)."""
    import json
    recs = [r for r in _traffic_lines() if "kernels" in r]
    assert recs
    for rec in recs:
        cnt = {"voxels": rec["voxels"], "bricks": rec["bricks"]}
        _, per = bench.step_bytes(cnt, rec["sweeps"])
        for kind, key in (("rhs", "rhs"), ("first_sweep", "first"), ("mid_sweep", "mid"),
                          ("last_sweep", "last")):
            measured = bench.profiled_traffic(cnt, kind)
            assert measured == rec["kernels"][kind]
            assert abs(measured / per[key] - 1.0) <= 0.05, (kind, measured, per[key])
