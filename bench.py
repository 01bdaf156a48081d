#!/usr/bin/env python3
"""Benchmark of the device-resident MPI hot path on one MI355X.

Workload (the p2p configuration): 256 MPI ranks, one workgroup each, in one
persistent kernel; every rank passes its 64 MiB seeded uint8 buffer to
rank+1 mod 256 (MPI_Irecv + MPI_Isend + MPI_Waitall).  A step is one ring
shift of all ranks.  value = whole-job throughput in GB/s with the
algorithmic byte count 2 * P * S per step (each payload read once and
written once), from HIP events around the timed launch; the in-kernel
MPI_Wtime timing (max over ranks) is reported beside it.  The received
buffers are verified against the seeded payloads after the timed launch.

Also reported on the same line: in-GPU ping-pong latency (8 B, one pair,
one-way = RTT/2) and the host-MPI baseline (our ring harness under MPICH on
the box's own cores, run before the GPU is touched).

usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import shutil
import subprocess
import sys

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "dphpc-gpu_mpi_amd"))

METRIC = "Allreduce/Alltoallv GB/s (% HBM peak); p2p latency µs; sort keys/s vs host MPI"
HBM_PEAK_GBPS = 8000.0
SEED = 0x5EED


def cpu_baseline(nbytes, budget_s=20.0):
    """Host MPI (MPICH) ring harness on the host cores; runs before any GPU use."""
    oracle = os.path.join(ROOT, "oracle")
    exe = os.path.join(oracle, "_build", "hostmpi_p2p")
    mpirun = "/opt/conda/bin/mpirun"
    if not os.path.exists(mpirun):
        return None
    if not os.path.exists(exe):
        subprocess.run(["make", "-s", "-C", oracle, "_build/hostmpi_p2p"], check=True,
                       stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
    cores = min(os.cpu_count() or 1, 16)
    # a short probe sizes the sample to roughly `budget_s` of work
    probe = subprocess.run([mpirun, "-np", str(cores), exe, "ring", str(nbytes), "2", "1", str(SEED)],
                           capture_output=True, text=True, timeout=300)
    p = json.loads(probe.stdout.strip().splitlines()[-1])
    per_iter = max(p["seconds"] / 2, 1e-6)
    iters = int(max(3, min(200, budget_s / per_iter)))
    out = subprocess.run([mpirun, "-np", str(cores), exe, "ring", str(nbytes), str(iters), "1", str(SEED)],
                         capture_output=True, text=True, timeout=600)
    r = json.loads(out.stdout.strip().splitlines()[-1])
    ver = subprocess.run([mpirun, "--version"], capture_output=True, text=True).stdout.split("\n")
    mpi_version = next((l.strip() for l in ver if "Version" in l), "MPICH")
    return {
        "value": round(r["gbps"], 3), "unit": "GB/s", "cores": cores, "kind": "port",
        "sample": (f"host MPI ring shift, {cores} ranks (one per core) x {nbytes} B, {iters} steps, "
                   f"same seeded payload and verification; {mpi_version}; mismatches={r['mismatches']}"),
    }


def measured_traffic(P, S, threads, steps):
    """HBM bytes per timed launch from the rocprofv3 FETCH_SIZE / WRITE_SIZE
    runs recorded in profiles/ (per ring step, calibrated against a plain
    copy kernel), when they were taken on this workload; else None."""
    path = os.path.join(ROOT, "profiles", "p2p_ring_traffic.json")
    if not os.path.exists(path):
        return None
    rec = json.load(open(path))
    if (rec.get("ranks"), rec.get("message_bytes"), rec.get("threads")) != (P, S, threads):
        return None
    return round(rec["hbm_bytes_per_step"] * steps)


def copy_calibration(nbytes):
    """Two launches of a plain dwordx4 copy of nbytes (read nbytes, write
    nbytes): what FETCH_SIZE / WRITE_SIZE read for it scales the ring's."""
    import gpumpi as gm
    gm.load()
    nbytes = nbytes // 16 * 16
    a = gm.DeviceBuffer(nbytes)
    b = gm.DeviceBuffer(nbytes)
    a.zero()
    for _ in range(2):
        gm.copy_calibration(b, a, nbytes)
    print(json.dumps({"copy_calibration_bytes": nbytes, "launches": 2}), flush=True)
    return 0


def side_measurements(gm, np, threads):
    """Same-run figures for the other headline rows: exact FP32 Allreduce
    (256 ranks x 16 MiB, algorithmic 2*P*N*4 bytes) and NPB IS class B on
    128 ranks (keys/s, and its Alltoallv key exchange, 2 * 2^25 * 4 bytes
    per iteration), with `threads` per rank as the ring."""
    out = {}
    P, n, iters = 256, 1 << 22, 10
    ib = gm.DeviceBuffer(4 * n * P)
    ob = gm.DeviceBuffer(4 * n * P)
    ib.zero()
    times = gm.DeviceBuffer(8 * P)
    gm.run("coll_reduce", P, [0, gm.DT["MPI_FLOAT"], gm.OP["MPI_SUM"], n, ib, ob, 4 * n, -1, 0, 1, iters, times, 0, 0],
           threads_per_rank=threads, timeout_s=300.0)
    t = float(times.download(np.float64).max()) / iters
    out["allreduce_f32_256x16MiB_gbps"] = round(2.0 * P * n * 4 / t / 1e9, 1)
    del ib, ob
    res = gm.DeviceBuffer(8 * 8)
    res.zero()
    tout = gm.DeviceBuffer(8 * 4)
    gm.run("npb_is", 128, [0, ord("B"), res, tout, 0, 10], threads_per_rank=threads, user_lds_bytes=4096 + 98304,
           heap_bytes=128 * (32 << 20), max_requests=64, timeout_s=120.0)
    r = res.download(np.int64)
    tt = tout.download(np.float64)
    out["npb_is_class_b_128_ranks"] = {
        "verified": bool(r[0] == 50 + 128), "mkeys_per_s": round(10 * (1 << 25) / tt[0] / 1e6, 1),
        "alltoallv_gbps": round(2 * 10 * (1 << 25) * 4 / tt[1] / 1e9, 1)}
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--ranks", type=int, default=256)
    ap.add_argument("--bytes", type=int, default=64 << 20)
    ap.add_argument("--threads", type=int, default=512,
                    help="threads per rank workgroup (512: eight waves, one rank per CU at 256 ranks)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extras", action="store_true", help="skip the Allreduce / NPB IS side measurements")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run the plain copy kernel over BYTES (reference for HBM counters) and exit")
    args = ap.parse_args(argv)
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (the whole MPI world lives in one device); "
              f"--gpus {args.gpus} is not supported", file=sys.stderr)
        return 2
    if args.copy_calibration:
        return copy_calibration(args.copy_calibration)

    cpu = None if args.no_cpu_baseline else cpu_baseline(args.bytes)

    import numpy as np
    import gpumpi as gm
    gm.load()
    P, S = args.ranks, args.bytes
    stride = (S + 4095) // 4096 * 4096
    sb = gm.DeviceBuffer(stride * P)
    rb = gm.DeviceBuffer(stride * P)
    rb.zero()
    gm.fill_payload(sb, stride, S, SEED, 0, P)

    def ring(program, iters):
        sz = gm.DeviceBuffer.from_array(np.array([S], dtype=np.uint64))
        it = gm.DeviceBuffer.from_array(np.array([iters], dtype=np.int64))
        wm = gm.DeviceBuffer.from_array(np.array([0], dtype=np.int64))
        times = gm.DeviceBuffer(8 * P)
        res = gm.run(program, P, [0, sz, 1, it, wm, sb, rb, stride, times],
                     threads_per_rank=args.threads, timeout_s=600.0)
        return res, times.download(np.float64)

    # Warm-up launch under its own kernel name (p2p_ring_warmup), then the
    # timed launch of exactly K steps (p2p_ring_sweep): a kernel trace of this
    # command lists the timed launch alone under p2p_ring_sweep.
    ring("p2p_ring_warmup", args.warmup)
    res, t = ring("p2p_ring_sweep", args.steps)
    kernel_s = res.kernel_ms / 1e3
    bytes_per_step = 2.0 * P * S
    gbps = bytes_per_step * args.steps / kernel_s / 1e9
    wtime_gbps = bytes_per_step * args.steps / float(t.max()) / 1e9
    bad = gm.verify_payload(rb, stride, S, SEED, -1, P, P)

    # in-GPU ping-pong latency, 8 B, pair (0, 1) inside the same 256-rank world,
    # with four waves per rank (idle helper waves poll LDS beside wave 0, so
    # fewer of them leave the small-message path less contended)
    ppb = gm.DeviceBuffer(4096 * P)
    sz = gm.DeviceBuffer.from_array(np.array([8], dtype=np.uint64))
    it = gm.DeviceBuffer.from_array(np.array([1000], dtype=np.int64))
    wm = gm.DeviceBuffer.from_array(np.array([100], dtype=np.int64))
    pt = gm.DeviceBuffer(8 * P)
    lat_threads = 256
    gm.run("p2p_pingpong_sweep", P, [0, sz, 1, it, wm, ppb, ppb, 4096, pt, 1],
           threads_per_rank=lat_threads, timeout_s=120.0)
    rtt_us = float(pt.download(np.float64)[:2].max()) / 1000 * 1e6

    line = {
        "metric": METRIC,
        "value": round(gbps, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(res.kernel_ms / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic: seeded splitmix64 uint8 payload, verified after the timed launch",
        "config": {"workload": f"p2p ring shift: {P} ranks (one workgroup each, {args.threads} threads) "
                               f"x {S} B per rank per step, MPI_Irecv+MPI_Isend+MPI_Waitall",
                   "ranks": P, "message_bytes": S, "parallelism": "1 GPU, 1 persistent kernel"},
        "roofline": {"bound": "hbm", "achieved": round(gbps, 2), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": round(gbps / HBM_PEAK_GBPS, 4),
                     "traffic": measured_traffic(P, S, args.threads, args.steps)},
        "cpu_baseline": cpu,
        "p2p_latency_us_8B_one_way": round(rtt_us / 2, 3),
        "p2p_rtt_us_8B": round(rtt_us, 3),
        "p2p_latency_threads_per_rank": lat_threads,
        "ring_gbps_mpi_wtime": round(wtime_gbps, 2),
        "payload_mismatches": bad,
    }
    if not args.no_extras:
        line.update(side_measurements(gm, np, args.threads))
    print(json.dumps(line), flush=True)
    del sb, rb, ppb
    gm.release_world()
    return 0 if bad == 0 else 1


if __name__ == "__main__":
    sys.exit(main())
