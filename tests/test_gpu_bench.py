"""The benchmark on the GPU, in-process: bench.py's main at its default
workload (256 ranks x 64 MiB ring, 1000 timed steps, verified payload),
checking the JSON line it prints, and the plain copy kernel the HBM traffic
counters are calibrated with.  These are also the commands the committed
rocprofv3 summaries were taken with (profiles/)."""
import json
import os
import sys

import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def test_bench_line(capsys):
    import bench
    rc = bench.main(["--no-cpu-baseline", "--no-extras"])
    line = json.loads(capsys.readouterr().out.strip().splitlines()[-1])
    print(json.dumps(line))
    assert rc == 0 and line["payload_mismatches"] == 0
    assert line["n_gpus"] == 1 and line["steps"] == 1000 and line["unit"] == "GB/s"
    assert 0 < line["roofline"]["frac"] < 1 and line["value"] > 0
    assert 0 < line["p2p_latency_us_8B_one_way"] < 100


def test_copy_calibration_kernel(capsys):
    import bench
    assert bench.copy_calibration(1 << 30) == 0
    assert json.loads(capsys.readouterr().out.strip())["copy_calibration_bytes"] == 1 << 30
