"""The driver's bench command, in a reduced form: `--steps 20 --warmup 5` on the batched
workload (4 members instead of 64, no shell legs, no served stream, a 1 s CPU sample).  The
result must be one JSON-serialisable record with the contract's keys, a roofline and a CPU
baseline (with its tight-tolerance and electrogram figures), and its validation members
within the LAT and electrogram bars."""
import json

import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.timeout(600)
def test_bench_driver_command_reduced(cuda):
    import bench
    rec = bench.run(["--gpus", "1", "--steps", "20", "--warmup", "5", "--sims", "4",
                     "--serve-requests", "0", "--cpu-seconds", "1", "--only-headline",
                     "--workers", "1"])
    line = json.loads(json.dumps(rec))
    for k in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step",
              "higher_is_better", "scaling", "vs_baseline", "dtype", "data", "config",
              "roofline", "cpu_baseline"):
        assert k in line, k
    assert line["steps"] == 20 and line["warmup"] == 5 and line["n_gpus"] == 1
    assert line["value"] > 0 and line["ms_per_step"] > 0
    assert abs(line["value"] - line["config"]["voxels"] / (line["ms_per_step"] * 1e-3)) \
        <= 1e-6 * line["value"]
    assert "workload" in line["config"]
    assert line["device"]["compute_units"] > 0 and "gfx950" in (line["device"]["arch"] or "")
    roof = line["roofline"]
    assert roof["bound"] == "hbm" and 0 < roof["frac"] < 1.5
    assert roof["achieved"] == pytest.approx(roof["frac"] * roof["peak"])
    cpu = line["cpu_baseline"]
    assert cpu["value"] > 0 and cpu["cores"] >= 1 and cpu["kind"] == "port"
    assert cpu["cpu_model"]
    assert cpu["tight"]["value"] > 0 and cpu["tight"]["steps"] > 0
    assert cpu["egm"]["value"] > 0 and cpu["egm"]["electrodes"] == 64
    for v in line["validate"]:
        assert v["lat_err_steps"]["max"] <= 1
        assert v["lat_err_steps"]["activation_mismatch"] == 0
        assert v["egm_rel_l2"] <= 1e-3
