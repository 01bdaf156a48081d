"""MPI_Wtime's scale on the device: the ring's timed region measured by
MPI_Wtime inside the kernel (s_memrealtime at the device's wall-clock rate)
against the HIP-event time of the whole launch on its stream.  The timed
region is most of the launch, so the in-kernel time is a large fraction of
the event time and never more than it (the reference's MPI_Wtime used
clock64() over the SM clock, mpi.cu:283-290, with no such check)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_wtime_agrees_with_the_launch_events(gm):
    P, S, iters = 16, 16 << 20, 40
    print(f"[gpu] wtime ring P={P} S={S} iters={iters}", flush=True)
    sb = gm.DeviceBuffer(S * P)
    rb = gm.DeviceBuffer(S * P)
    gm.fill_payload(sb, S, S, 11, 0, P)
    sz = gm.DeviceBuffer.from_array(np.array([S], dtype=np.uint64))
    it = gm.DeviceBuffer.from_array(np.array([iters], dtype=np.int64))
    wm = gm.DeviceBuffer.from_array(np.array([2], dtype=np.int64))
    times = gm.DeviceBuffer(8 * P)
    res = gm.run("p2p_ring_sweep", P, [0, sz, 1, it, wm, sb, rb, S, times], timeout_s=60.0)
    assert res.status == 0, res.describe()
    wtime = float(times.download(np.float64).max())
    events = res.kernel_ms / 1e3
    print(f"[gpu] in-kernel {wtime * 1e3:.3f} ms, events {events * 1e3:.3f} ms", flush=True)
    assert 0.5 * events <= wtime <= 1.02 * events
