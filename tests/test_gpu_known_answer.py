"""The reference's GPU known-answer tests, run as rank programs on the HIP
runtime (tests/cuda_mpi_test.cu, collectives_test.cu, communicator_test.cu;
rank counts as there)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CASES = [
    ("ka_single_int", 2, {}),
    ("ka_array16", 2, {}),
    ("ka_sendrecv_pm", 2, {}),
    ("ka_repeat_tags", 2, {}),
    ("ka_network_flood", 3, {}),
    ("ka_all_to_all_fifo", 10, {"max_requests": 400}),
    ("ka_reduce_prod_i16", 5, {}),
    ("ka_allreduce_double", 7, {}),
    ("ka_allgather", 7, {}),
    ("ka_scatter", 5, {}),
    ("ka_scatterv", 5, {}),
    ("ka_comm_rank", 5, {}),
    ("ka_comm_dup_isolation", 2, {}),
    ("ka_comm_split", 10, {}),
]


@pytest.mark.parametrize("threads", [64, 256])
@pytest.mark.parametrize("name,nranks,kw", CASES, ids=[c[0] for c in CASES])
def test_known_answer(gm, name, nranks, kw, threads):
    print(f"[gpu] {name} nranks={nranks} threads={threads}", flush=True)
    ok = gm.DeviceBuffer(4 * nranks)
    ok.zero()
    res = gm.run(name, nranks, [ok], threads_per_rank=threads, timeout_s=30.0, **kw)
    assert res.status == 0, res.describe()
    got = ok.download(np.int32)
    assert got.tolist() == [1] * nranks, got
