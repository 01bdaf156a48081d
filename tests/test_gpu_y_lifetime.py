"""Resource lifetime within a launch and user-defined reduction operators
on the device runtime (progs/lifetime.hip, progs/user_ops.hip):
  * the rank heap reuses, splits and merges freed blocks and lowers its top
    (exact block positions), then survives random churn that a bump
    allocator would exhaust;
  * MPI_Comm_free returns communicator blocks and context ids: more
    creations than there are context ids (32765) and than a 64 KiB pool
    holds, every one used by an Allreduce;
  * MPI_Comm_free completes requests still pending on the communicator and
    their status names the sender's rank in it;
  * MPI_Op_create operators in MPI_Reduce / MPI_Allreduce: a user FP64 sum
    equals MPI_SUM bit for bit, an asymmetric int64 operator equals the
    closed form of the reference's fold order (mpi.cu:209-225)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def run_lifetime(gm, mode, P, iters, **kw):
    ok = gm.DeviceBuffer(4 * P)
    ok.zero()
    out = gm.DeviceBuffer(8 * 4)
    out.zero()
    res = gm.run("lifetime", P, [ok, mode, iters, out], timeout_s=120.0, **kw)
    assert res.status == 0, res.describe()
    return ok.download(np.int32), out.download(np.int64)


@pytest.mark.parametrize("P", [1, 4, 64])
def test_rank_heap_reuse(gm, P):
    print(f"[gpu] lifetime heap P={P}", flush=True)
    ok, out = run_lifetime(gm, 0, P, 20000, heap_bytes=P * (2 << 20))
    assert ok.tolist() == [1] * P and out[0] == 20000


def test_comm_free_returns_context_ids_and_pool_blocks(gm):
    print("[gpu] lifetime communicator churn P=4, 33000 iterations", flush=True)
    ok, out = run_lifetime(gm, 1, 4, 33000, comm_pool_bytes=64 << 10)
    assert ok.tolist() == [1] * 4 and out[0] == 33000


@pytest.mark.parametrize("P", [2, 7])
def test_comm_free_completes_pending_requests(gm, P):
    print(f"[gpu] lifetime pending requests P={P}", flush=True)
    ok, _ = run_lifetime(gm, 2, P, 0)
    assert ok.tolist() == [1] * P


def _expected_twice_plus(P, n):
    i = np.arange(n, dtype=np.int64)
    x = lambda r: (r * 7919 + i * 104729) % 100003 - 50000  # noqa: E731
    return x(0) + 2 * sum((x(r) for r in range(1, P)), np.zeros(n, dtype=np.int64))


@pytest.mark.parametrize("P,n,root,in_place", [(1, 10, -1, 1), (2, 1, 1, 0), (3, 1000, -1, 0), (4, 63, 2, 1),
                                               (5, 4097, 0, 0), (16, 100003, -1, 1), (64, 70001, 17, 0)])
def test_user_defined_operators(gm, P, n, root, in_place):
    print(f"[gpu] user_op_reduce P={P} n={n} root={root} in_place={in_place}", flush=True)
    ok = gm.DeviceBuffer(4 * P)
    ok.zero()
    out = gm.DeviceBuffer(8 * n)
    out.zero()
    res = gm.run("user_op_reduce", P, [ok, n, root, in_place, out], heap_bytes=P * (8 << 20), timeout_s=120.0)
    assert res.status == 0, res.describe()
    assert ok.download(np.int32).tolist() == [1] * P
    assert np.array_equal(out.download(np.int64), _expected_twice_plus(P, n))
