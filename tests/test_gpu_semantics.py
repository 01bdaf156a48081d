"""Point-to-point, request and communicator semantics on the GPU against the
same scenarios run under host MPICH (tests/golden/semantics_mpich.json).

The scenarios (dphpc-gpu_mpi_amd/csrc/progs/semantics_scenarios.h) cover
MPI_Sendrecv, MPI_ANY_SOURCE / MPI_ANY_TAG receives of queued and of later
messages, the matching order of posted receives, MPI_Test / Testall / Waitany
/ Waitsome, a non-blocking MPI_Request_free, MPI_Ssend / Issend, MPI_Status
with MPI_Get_count, MPI_PROC_NULL, and Group_incl / Group_range_incl /
Comm_create / Comm_split(MPI_UNDEFINED).  Every rank's record must equal
host MPI's exactly."""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "semantics_mpich.json")
CAP = 512
NAMES = ["sendrecv", "any_source", "any_tag", "posted_order", "completion", "request_free", "sync_status",
         "groups"]


def _golden(P):
    return json.load(open(GOLDEN))["runs"][str(P)]


def _run(gm, P, mask, threads):
    rec = gm.DeviceBuffer(4 * CAP * P)
    rec.zero()
    lens = gm.DeviceBuffer(4 * P)
    lens.zero()
    res = gm.run("semantics", P, [rec, CAP, mask, lens], threads_per_rank=threads, timeout_s=60.0)
    assert res.status == 0, res.describe()
    n = lens.download(np.int32)
    v = rec.download(np.int32).reshape(P, CAP)
    return [v[r, :n[r]].tolist() for r in range(P)]


@pytest.mark.parametrize("threads", [64, 256])
@pytest.mark.parametrize("P", [4, 8, 13])
def test_semantics_match_host_mpi(gm, P, threads):
    gold = _golden(P)
    print(f"[gpu] semantics P={P} threads={threads}", flush=True)
    got = _run(gm, P, gold["mask"], threads)
    for r in range(P):
        assert got[r] == gold["records"][r], f"rank {r}: {got[r]} != {gold['records'][r]}"


@pytest.mark.parametrize("k", range(len(NAMES)), ids=NAMES)
def test_each_scenario_alone(gm, k):
    """Each scenario on its own at P=8, against host MPI's run of the same
    scenario alone: a failure names the scenario."""
    P = 8
    print(f"[gpu] semantics scenario {NAMES[k]}", flush=True)
    got = _run(gm, P, 1 << k, 256)
    gold = _golden(P)["alone"][k]
    for r in range(P):
        assert got[r] == gold[r], f"rank {r}: {got[r]} != {gold[r]}"


@pytest.mark.parametrize("rendezvous", [0, 1])
def test_truncated_receive_is_reported(gm, rendezvous):
    """A receive buffer smaller than the message: the launch ends with
    MPI_ERR_TRUNCATE recorded for the receiving rank (errors are fatal, as
    under MPI_ERRORS_ARE_FATAL)."""
    print(f"[gpu] sem_truncate rendezvous={rendezvous}", flush=True)
    res = gm.run("sem_truncate", 2, [0, rendezvous], check=False, timeout_s=30.0)
    assert res.status == 4, res.describe()  # EDEVICE
    assert res.mpi_error == 20 and res.mpi_error_rank == 0, res.describe()
