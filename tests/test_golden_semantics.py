"""The semantics fixture (tests/golden/semantics_mpich.json) is what host MPI
gives for the scenarios: re-run them under MPICH where it is installed and
compare every rank's record with the committed one (pins the fixture the GPU
tests compare with)."""
import json
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
MPIRUN = "/opt/conda/bin/mpirun"
EXE = os.path.join(ROOT, "oracle", "_build", "hostmpi_semantics_golden")
GOLDEN = os.path.join(HERE, "golden", "semantics_mpich.json")


@pytest.fixture(scope="module")
def exe():
    if not os.path.exists(MPIRUN):
        pytest.skip("host MPI (MPICH at /opt/conda/bin) is not installed here")
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "_build/hostmpi_semantics_golden"],
                   check=True, capture_output=True)
    return EXE


@pytest.mark.parametrize("P", [4, 8])
def test_fixture_is_host_mpi_output(exe, P):
    gold = json.load(open(GOLDEN))["runs"][str(P)]
    out = subprocess.run([MPIRUN, "-np", str(P), exe], check=True, capture_output=True, text=True,
                         timeout=120).stdout
    assert json.loads(out)["records"] == gold["records"]


def test_fixture_shape():
    runs = json.load(open(GOLDEN))["runs"]
    for P, run in runs.items():
        P = int(P)
        assert run["P"] == P and len(run["records"]) == P and len(run["alone"]) == 8
        # Scenario records concatenate to the full run's record.
        for r in range(P):
            assert sum((a[r] for a in run["alone"]), []) == run["records"][r]
        # MPI-fixed facts, independent of any implementation:
        rec0 = run["records"][0]
        # posted-receive order (scenario 4): a=71, b=72, c=81, d=82
        assert [71, 72, 81, 82] == run["alone"][3][0][:4]
        # ANY_SOURCE per-source order, queued and posted (scenario 2)
        exp = [s * 100 + j for s in range(1, P) for j in range(3)]
        alone = run["alone"][1][0]
        assert alone[0] == 0 and alone[1:1 + len(exp)] == exp
        assert alone[1 + len(exp)] == 0 and alone[2 + len(exp):] == exp
        assert rec0  # non-empty
