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


def test_profiled_traffic_matches_measured_arena_only():
    cnt = {"voxels": 3191136, "bricks": 81904}
    t = bench.profiled_traffic(cnt)
    assert isinstance(t, int) and 60e6 < t < 100e6
    assert bench.profiled_traffic({"voxels": 1000, "bricks": 20}) is None


def test_profiled_traffic_of_the_electrogram():
    rec = bench._profile_record(bench.PROFILE_TRAFFIC_EGM,
                                {"n_voxel": 199446, "electrodes": 64, "samples": 2000},
                                "bytes_per_launch")
    algorithmic = 4.0 * 2000 * 199446 + 4.0 * 64 * 199446 + 4.0 * 2000 * 64
    assert rec is not None and algorithmic < rec < 1.3 * algorithmic
    assert bench._profile_record(bench.PROFILE_TRAFFIC_EGM,
                                 {"n_voxel": 199446, "electrodes": 256, "samples": 2000},
                                 "bytes_per_launch") is None
