"""Regenerates the host-MPI golden fixtures in this directory.

Runs in the build container (it needs host MPICH at /opt/conda/bin and, for
the pi fixture, the reference's cpi.c under /root/reference):

  * reduce_mpich.bin -- oracle/hostmpi/reduce_golden.c under `mpirun -np 5`:
    rank 0's MPI_Reduce result for every legal (datatype, op) pair.
  * cpi_mpich.json   -- the reference's toolchain_tests/pi/cpi.c compiled with
    host mpicc (oracle/Makefile `ref`) and run with N=1e8 at several rank
    counts; the printed pi is recorded.
  * ring_mpich.json -- dphpc-gpu_mpi_amd/examples/ring.c (the source
    converter's example C program) compiled with host mpicc and run with 5
    rounds at 1, 4, 7 and 64 ranks: rank 0's output line.
  * semantics_mpich.json -- the point-to-point / request / communicator
    scenarios (dphpc-gpu_mpi_amd/csrc/progs/semantics_scenarios.h) under
    host MPI (oracle/hostmpi/semantics_golden.c) at 4, 8 and 13 ranks: every
    rank's record.
  * npb_dt_mpich.json -- the reference's NPB DT class S (oracle/Makefile
    `ref`) under host MPI for the BH and WH graphs at 5 ranks and SH at 12:
    the printed L2 norm and NPB's verification.
"""
import json
import os
import re
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
ORACLE = os.path.join(ROOT, "oracle")
MPIRUN = "/opt/conda/bin/mpirun"


def run(cmd, **kw):
    print("+", " ".join(cmd), flush=True)
    return subprocess.run(cmd, check=True, **kw)


def semantics():
    run(["make", "-C", ORACLE, "_build/hostmpi_semantics_golden"])
    exe = os.path.join(ORACLE, "_build", "hostmpi_semantics_golden")
    out = {}
    for np_ in (4, 8, 13):
        txt = run([MPIRUN, "-np", str(np_), exe], capture_output=True, text=True).stdout
        out[str(np_)] = json.loads(txt)
        # Each scenario alone, so a GPU failure names its scenario.
        out[str(np_)]["alone"] = []
        for k in range(8):
            txt = run([MPIRUN, "-np", str(np_), exe, str(1 << k)], capture_output=True, text=True).stdout
            out[str(np_)]["alone"].append(json.loads(txt)["records"])
    with open(os.path.join(HERE, "semantics_mpich.json"), "w") as f:
        json.dump({"source": "semantics_scenarios.h under MPICH 3.3.2", "runs": out}, f)


def ring():
    src = os.path.join(ROOT, "dphpc-gpu_mpi_amd", "examples", "ring.c")
    exe = os.path.join(ORACLE, "_build", "hostmpi_ring")
    os.makedirs(os.path.dirname(exe), exist_ok=True)
    run(["/opt/conda/bin/mpicc", "-cc=gcc", "-O2", "-o", exe, src])
    out = {}
    for np_ in (1, 4, 7, 64):
        txt = run([MPIRUN, "-np", str(np_), exe, "5"], capture_output=True, text=True).stdout
        out[str(np_)] = [l for l in txt.splitlines() if l.startswith("ring:")][0]
    with open(os.path.join(HERE, "ring_mpich.json"), "w") as f:
        json.dump(out, f, indent=1)


def npb_dt():
    run(["make", "-s", "-C", ORACLE, "_ref/npb/dt.S.x"])
    exe = os.path.join(ORACLE, "_ref", "npb", "dt.S.x")
    out = {}
    for graph, np_ in (("BH", 5), ("WH", 5), ("SH", 12)):
        txt = run([MPIRUN, "-np", str(np_), exe, graph], stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                  text=True).stdout  # the norm is printed to stderr
        norm = re.search(r"DT_\w+\.S L2 Norm = ([0-9.]+)", txt).group(1)
        ok = re.search(r"Verification\s+=\s+SUCCESSFUL", txt) is not None
        out[graph] = {"ranks": np_, "l2_norm": norm, "verified": ok}
    with open(os.path.join(HERE, "npb_dt_mpich.json"), "w") as f:
        json.dump({"source": "the reference's NPB DT class S under MPICH 3.3.2", "runs": out}, f, indent=1)


def main():
    if len(sys.argv) > 1 and sys.argv[1] == "semantics":
        semantics()
        return 0
    if len(sys.argv) > 1 and sys.argv[1] == "ring":
        ring()
        return 0
    if len(sys.argv) > 1 and sys.argv[1] == "npb_dt":
        npb_dt()
        return 0
    semantics()
    ring()
    npb_dt()
    run(["make", "-C", ORACLE, "_build/hostmpi_reduce_golden", "ref"])
    run([MPIRUN, "-np", "5", os.path.join(ORACLE, "_build", "hostmpi_reduce_golden"),
         os.path.join(HERE, "reduce_mpich.bin")])
    pis = {}
    for np_ in (1, 8, 64):
        out = run([MPIRUN, "-np", str(np_), os.path.join(ORACLE, "_ref", "cpi"), "100000000"],
                  capture_output=True, text=True).stdout
        m = re.search(r"pi is approximately ([0-9.]+)", out)
        pis[str(np_)] = m.group(1)
    with open(os.path.join(HERE, "cpi_mpich.json"), "w") as f:
        json.dump({"n": 100000000, "pi_printed": pis,
                   "source": "toolchain_tests/pi/cpi.c under MPICH 3.3.2"}, f, indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main())
