"""bench.py's command line and its side-configuration plumbing, on the CPU
(no GPU here): more than one GPU is refused with a message, every side
configuration runs in a child process and a child that fails leaves a
`status` entry instead of ending the run, and the CPU baselines the line
carries are produced in the form the line expects."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_more_than_one_gpu_is_refused():
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert p.returncode == 2 and "one GPU" in p.stderr and not p.stdout.strip()


def test_side_configurations_are_isolated(monkeypatch):
    # Without a GPU every child fails; each failure is recorded under its own
    # name and the others still run.
    import argparse
    args = argparse.Namespace(no_cpu_baseline=True, wire=False, threads=256)
    monkeypatch.setattr(bench, "CONFIGS", ("pi", "allreduce"))
    out = bench.side_configs(args)
    assert set(out) == {"pi", "allreduce"}
    for v in out.values():
        assert v["status"].startswith("failed: exit status"), v
    json.dumps(out)  # the line stays JSON


@pytest.mark.skipif(not os.path.exists("/opt/conda/bin/mpirun"), reason="host MPI is absent")
def test_is_host_baseline_shape():
    r = bench.is_host_baseline()
    assert r["unit"] == "Mkeys/s" and r["value"] > 0 and r["kind"] in ("reference", "port")
    assert r["cores"] >= 1 and r["sample"]
