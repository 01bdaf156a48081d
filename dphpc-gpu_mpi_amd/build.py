"""Builds the runtime's shared libraries in-tree for gfx950.

  lib/libgpumpi.so        host C-ABI (include/gpumpi/host.h): launcher, watchdog
  lib/libgpumpi_progs.so  rank programs (tests, workloads, benchmarks) with a
                          by-name registry, linked against libgpumpi.so
  lib/ring, lib/cpi       C MPI programs converted by converter.py and built by
                          gpumpicc.py (cpi from the reference's source, when present)

Usage: python dphpc-gpu_mpi_amd/build.py [-j N] [--force]
"""
import argparse
import concurrent.futures as cf
import glob
import os
import shutil
import subprocess
import sys
import time

PKG = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, PKG)
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "lib")
OBJ = os.path.join(PKG, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
REFERENCE = os.environ.get("GPUMPI_REFERENCE", "/root/reference")

COMMON = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"), "-I", CSRC,
    "-Wno-unused-result",
]


def _headers():
    return glob.glob(os.path.join(ROOT, "include", "**", "*.h"), recursive=True) + \
        glob.glob(os.path.join(CSRC, "**", "*.hpp"), recursive=True)


def _stale(target, deps):
    if not os.path.exists(target):
        return True
    t = os.path.getmtime(target)
    return any(os.path.getmtime(d) > t for d in deps)


def _compile(src, obj, force):
    deps = [src] + _headers()
    if not force and not _stale(obj, deps):
        return obj, "", 0.0
    import time
    t0 = time.time()
    cmd = [HIPCC] + COMMON + ["-c", src, "-o", obj]
    p = subprocess.run(cmd, capture_output=True, text=True)
    err = "\n".join(l for l in (p.stdout + p.stderr).splitlines()
                    if "packed-fp32-ops" not in l)
    if p.returncode != 0:
        raise RuntimeError(f"compile failed: {src}\n{err}")
    return obj, err, time.time() - t0


def build(jobs=8, force=False, verbose=True):
    os.makedirs(LIB, exist_ok=True)
    os.makedirs(OBJ, exist_ok=True)
    host_src = [os.path.join(CSRC, "host", "launcher.cpp"), os.path.join(CSRC, "host", "service.cpp"),
                os.path.join(CSRC, "host", "noalloc.hip")]
    prog_src = sorted(glob.glob(os.path.join(CSRC, "progs", "*.hip"))) + \
        [os.path.join(CSRC, "progs", "registry.cpp")]
    units = [(s, os.path.join(OBJ, os.path.basename(s) + ".o")) for s in host_src + prog_src]
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        futs = {ex.submit(_compile, s, o, force): s for s, o in units}
        for f in cf.as_completed(futs):
            obj, msg, dt = f.result()
            if verbose and dt:
                print(f"[build] {os.path.basename(futs[f])} {dt:.1f}s", flush=True)
            if verbose and msg.strip():
                print(msg, flush=True)
    core = os.path.join(LIB, "libgpumpi.so")
    host_objs = [o for s, o in units if s in host_src]
    if force or _stale(core, host_objs):
        subprocess.run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", core] + host_objs,
                       check=True)
    progs = os.path.join(LIB, "libgpumpi_progs.so")
    prog_objs = [o for s, o in units if s in prog_src]
    if force or _stale(progs, prog_objs + [core]):
        subprocess.run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", progs] + prog_objs +
                       ["-L", LIB, "-lgpumpi", "-Wl,-rpath,$ORIGIN"], check=True)
    # Converted C programs (converter.py + gpumpicc.py): standalone
    # executables with the reference's `---gpumpi` command line.  lib/ring is
    # this package's example; lib/cpi is the reference's own pi program,
    # converted from its source when that is at hand.
    # lib/is.S.x is the reference's NPB IS (is.c with
    # common/c_print_results.c and common/c_timers.c, unmodified), lib/dt.S.x
    # its NPB DT (dt.c, DGraph.c and the common files, unmodified).
    import gpumpicc
    examples = [([os.path.join(PKG, "examples", "ring.c")], os.path.join(LIB, "ring"), [])]
    ref_cpi = os.path.join(REFERENCE, "toolchain_tests", "pi", "cpi.c")
    if os.path.exists(ref_cpi):
        examples.append(([ref_cpi], os.path.join(LIB, "cpi"), []))
    # the reference's own converter tests: globals (toolchain_tests/global_var),
    # MPI-IO (basicIO/basicIO.c, basicIO/file_view1.c) and a contiguous
    # datatype (datatype/contiguous.c)
    for rel, exe in (("global_var/global_var.c", "global_var"), ("basicIO/basicIO.c", "basicIO"),
                     ("datatype/contiguous.c", "contiguous"), ("basicIO/file_view1.c", "file_view1")):
        src = os.path.join(REFERENCE, "toolchain_tests", rel)
        if os.path.exists(src):
            examples.append(([src], os.path.join(LIB, exe), []))
    if os.path.exists(os.path.join(NPB, "IS", "is.c")) and shutil.which("gcc"):
        for cls in ("S",):  # the program runs as scalar code on each rank's wave 0: class S
            srcs = [os.path.join(NPB, "IS", "is.c"), os.path.join(NPB, "common", "c_print_results.c"),
                    os.path.join(NPB, "common", "c_timers.c")]
            examples.append((srcs, os.path.join(LIB, f"is.{cls}.x"), [npb_params("is", cls)]))
        dt = [os.path.join(NPB, "DT", "dt.c"), os.path.join(NPB, "DT", "DGraph.c")] + \
            [os.path.join(NPB, "common", c) for c in ("c_print_results.c", "c_timers.c", "randdp.c")]
        examples.append((dt, os.path.join(LIB, "dt.S.x"), [npb_params("dt", "S")]))
    conv_deps = [core, os.path.join(PKG, "converter.py"), os.path.join(PKG, "gpumpicc.py")] + _headers()

    def convert_one(srcs, exe, incs):
        if not (force or _stale(exe, srcs + conv_deps)):
            return
        t0 = time.time()
        work = os.path.join(OBJ, "conv_" + os.path.basename(exe))
        os.makedirs(work, exist_ok=True)
        gpumpicc.build(srcs, exe, include_dirs=incs, workdir=work, rpath="$ORIGIN")
        if verbose:
            print(f"[build] {os.path.basename(exe)} (converted {', '.join(os.path.basename(x) for x in srcs)}) "
                  f"{time.time() - t0:.1f}s", flush=True)

    with cf.ThreadPoolExecutor(max_workers=min(jobs, len(examples))) as ex:
        for f in [ex.submit(convert_one, *e) for e in examples]:
            f.result()
    return core, progs


NPB = os.path.join(REFERENCE, "toolchain_tests", "npb", "NPB3.4.1", "NPB3.4-MPI")


def npb_params(bench, cls):
    """Directory holding npbparams.h of NPB benchmark `bench`, class `cls`,
    made by the reference's own procedure (toolchain_tests/npb/is_test.sh:
    11-15): config/make.def is its make.def.template with MPICC set to this
    package's gpumpicc, and its sys/setparams.c writes the header."""
    d = os.path.join(OBJ, "npb")
    wd = os.path.join(d, f"{bench}_{cls}")
    if os.path.exists(os.path.join(wd, "npbparams.h")):
        return wd  # made already (it depends on the class alone)
    os.makedirs(os.path.join(d, "config"), exist_ok=True)
    with open(os.path.join(NPB, "config", "make.def.template")) as f:
        tmpl = f.read()
    with open(os.path.join(d, "config", "make.def"), "w") as f:
        f.write(tmpl.replace("MPICC = mpicc", "MPICC = " + os.path.join(PKG, "gpumpicc.py")))
    sp = os.path.join(d, "setparams")
    if not os.path.exists(sp):
        subprocess.run(["gcc", "-O2", "-o", sp, os.path.join(NPB, "sys", "setparams.c")], check=True,
                       capture_output=True)
    os.makedirs(wd, exist_ok=True)
    subprocess.run([sp, bench, cls], cwd=wd, check=True, stdout=subprocess.DEVNULL)
    return wd


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("-j", type=int, default=8)
    ap.add_argument("--force", action="store_true")
    a = ap.parse_args()
    build(a.j, a.force)
    return 0


if __name__ == "__main__":
    sys.exit(main())
