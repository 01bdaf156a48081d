"""gpumpi_run_noalloc -- the launch with its whole world in the library's
static device array (no allocation) -- gives the same results as gpumpi_run,
on two calls in a row, the second with a smaller input; and a world that
does not fit is refused before anything runs."""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
SMALL = dict(staging_bytes=256 << 10, comm_pool_bytes=1 << 20, max_requests=64)
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "semantics_mpich.json")


def _allreduce(gm, x, noalloc):
    P, n = x.shape
    ib = gm.DeviceBuffer.from_array(x)
    ob = gm.DeviceBuffer(8 * n * P)
    ob.zero()
    res = gm.run("coll_reduce", P, [0, gm.DT["MPI_DOUBLE"], gm.OP["MPI_SUM"], n, ib, ob, 8 * n, -1, 0, 1, 1, 0, 0],
                 noalloc=noalloc, heap_bytes=P << 20, timeout_s=60.0, **SMALL)
    assert res.status == 0, res.describe()
    return ob.download(np.float64).reshape(P, n)


def test_noalloc_allreduce_matches_run_twice(gm, orc):
    rng = np.random.default_rng(31)
    for P, n in ((8, 50000), (6, 777)):  # second call: a smaller input
        x = rng.standard_normal((P, n))
        print(f"[gpu] noalloc allreduce P={P} n={n}", flush=True)
        got = _allreduce(gm, x, True)
        ref = _allreduce(gm, x, False)
        assert got.tobytes() == ref.tobytes()
        exp = orc.reduce_fold("MPI_DOUBLE", "MPI_SUM", x)
        for r in range(P):
            assert got[r].tobytes() == exp.tobytes(), r


def test_noalloc_semantics_match_run(gm):
    """p2p, requests and communicators in the static world: the semantics
    records equal gpumpi_run's and host MPI's."""
    cap = 512
    for P in (8, 4):
        recs = []
        for noalloc in (True, False):
            rec = gm.DeviceBuffer(4 * cap * P)
            rec.zero()
            lens = gm.DeviceBuffer(4 * P)
            lens.zero()
            print(f"[gpu] noalloc semantics P={P} noalloc={noalloc}", flush=True)
            res = gm.run("semantics", P, [rec, cap, 255, lens], noalloc=noalloc, heap_bytes=P << 20,
                         timeout_s=60.0, **SMALL)
            assert res.status == 0, res.describe()
            n = lens.download(np.int32)
            v = rec.download(np.int32).reshape(P, cap)
            recs.append([v[r, :n[r]].tolist() for r in range(P)])
        assert recs[0] == recs[1]
        assert recs[0] == json.load(open(GOLDEN))["runs"][str(P)]["records"]


def test_noalloc_refuses_a_world_that_does_not_fit(gm):
    print("[gpu] noalloc too-large world (refused on the host)", flush=True)
    res = gm.run("coll_reduce", 64, [0] * 14, check=False, noalloc=True)  # default 4 MiB staging/heap per rank
    assert res.status == 1  # GPUMPI_EINVAL, nothing launched
    assert "static world holds" in gm.last_error()
