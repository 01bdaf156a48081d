"""Point-to-point parity: ring shift and ping-pong payloads delivered by the
HIP runtime equal the oracle's seeded payloads byte for byte (small sizes,
downloaded and compared on the CPU), and the full-size runs keep the
size-independent property recv(r) == payload(r - 1) (checked on device by
regenerating the payload)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SEED = 0x5EED


def _ring(gm, nranks, sizes, iters=3, warm=1, threads=256, **kw):
    stride = (max(sizes) + 255) // 256 * 256
    sb = gm.DeviceBuffer(stride * nranks)
    rb = gm.DeviceBuffer(stride * nranks)
    rb.zero()
    gm.fill_payload(sb, stride, stride, SEED, 0, nranks)
    sz = gm.DeviceBuffer.from_array(np.array(sizes, dtype=np.uint64))
    it = gm.DeviceBuffer.from_array(np.full(len(sizes), iters, dtype=np.int64))
    wm = gm.DeviceBuffer.from_array(np.full(len(sizes), warm, dtype=np.int64))
    times = gm.DeviceBuffer(8 * len(sizes) * nranks)
    res = gm.run("p2p_ring_sweep", nranks, [0, sz, len(sizes), it, wm, sb, rb, stride, times],
                 threads_per_rank=threads, timeout_s=60.0, **kw)
    return res, sb, rb, stride, times


@pytest.mark.parametrize("threads", [64, 256])
def test_ring_payload_matches_oracle(gm, orc, threads):
    P = 8
    # eager/rendezvous boundary (448), unaligned sizes, helper-wave sizes
    sizes = [1, 8, 447, 448, 449, 4096, 65536 + 3, (1 << 20) + 5]
    for s in sizes:
        print(f"[gpu] ring P={P} size={s} threads={threads}", flush=True)
        res, sb, rb, stride, _ = _ring(gm, P, [s], threads=threads)
        assert res.status == 0, res.describe()
        for r in range(P):
            got = rb.download(np.uint8, s, offset=r * stride)
            exp = orc.payload_bytes(SEED, (r - 1) % P, stride)[:s]
            assert np.array_equal(got, exp), (s, r)


def test_ring_sweep_many_ranks_device_verified(gm):
    P = 256
    sizes = [8, 1 << 10, 1 << 16, 1 << 20]
    print(f"[gpu] ring sweep P={P}", flush=True)
    res, sb, rb, stride, times = _ring(gm, P, sizes, iters=5)
    assert res.status == 0, res.describe()
    # the last size's data remains in the receive buffers
    assert gm.verify_payload(rb, stride, sizes[-1], SEED, -1, P, P) == 0
    t = times.download(np.float64).reshape(len(sizes), P)
    assert (t > 0).all()


def test_ring_64mib_device_verified(gm):
    P = 32
    s = 64 << 20
    print(f"[gpu] ring P={P} size=64MiB", flush=True)
    res, sb, rb, stride, _ = _ring(gm, P, [s], iters=2)
    assert res.status == 0, res.describe()
    assert gm.verify_payload(rb, stride, s, SEED, -1, P, P) == 0
    assert gm.verify_payload(sb, stride, s, SEED, 0, P, P) == 0  # sources untouched


@pytest.mark.parametrize("s", [8, 448, 449, 100000])
def test_pingpong_payload(gm, orc, s):
    P = 4
    stride = (s + 255) // 256 * 256
    sb = gm.DeviceBuffer(stride * P)
    rb = gm.DeviceBuffer(stride * P)
    rb.zero()
    gm.fill_payload(sb, stride, stride, SEED, 0, P)
    sz = gm.DeviceBuffer.from_array(np.array([s], dtype=np.uint64))
    it = gm.DeviceBuffer.from_array(np.array([10], dtype=np.int64))
    wm = gm.DeviceBuffer.from_array(np.array([2], dtype=np.int64))
    times = gm.DeviceBuffer(8 * P)
    print(f"[gpu] pingpong P={P} size={s}", flush=True)
    res = gm.run("p2p_pingpong_sweep", P, [0, sz, 1, it, wm, sb, rb, stride, times, 2], timeout_s=60.0)
    assert res.status == 0, res.describe()
    for r in range(P):
        got = rb.download(np.uint8, s, offset=r * stride)
        assert np.array_equal(got, orc.payload_bytes(SEED, r ^ 1, stride)[:s]), r
    t = times.download(np.float64)
    assert (t > 0).all()
