"""gpumpicc: compile C MPI programs into executables of this runtime.

Follows the reference's scripts/gpumpicc.py:12-201 (sort the command line
into sources and compiler arguments, convert each source, compile and link
against the runtime): each C source goes through converter.py into HIP
source, then hipcc builds it for gfx950 with the runtime's headers and links
libgpumpi.so.  The executable takes the program's own arguments followed by
`---gpumpi -n RANKS ...` (include/gpumpi/host.h, gpumpi_parse_args).

    python dphpc-gpu_mpi_amd/gpumpicc.py prog.c [more.c ...] -o prog [-I dir] [-D name] [-O2] [-lm]

Several sources are compiled with relocatable device code (-fgpu-rdc), so
their functions and per-rank globals link across files.
"""
import os
import subprocess
import sys
import tempfile

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
sys.path.insert(0, PKG)
import converter  # noqa: E402

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
SOURCE_EXT = (".c", ".cpp", ".cxx")
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-ffp-contract=off",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wno-unused-result",
         "-Wno-format-security", "-Wno-unused-value", "-Wno-macro-redefined"]


def split_args(argv):
    """(sources, include dirs, hipcc args, output)."""
    sources, incs, args, out = [], [], [], "a.out"
    it = iter(argv)
    for a in it:
        if a.lower().endswith(SOURCE_EXT):
            sources.append(a)
        elif a == "-o":
            out = next(it)
        elif a in ("-I", "-D", "-L", "-l"):
            v = next(it)
            (incs if a == "-I" else args).extend([v] if a == "-I" else [a + v])
        elif a.startswith("-I"):
            incs.append(a[2:])
        elif a.startswith(("-D", "-L", "-l", "-O", "-g")) or a.endswith(".o"):
            args.append(a)
        else:
            print(f"gpumpicc: ignoring {a}", file=sys.stderr)
    return sources, incs, args, out


def build(sources, output, include_dirs=(), extra=(), workdir=None, verbose=False, rpath=None):
    """Convert and compile `sources` into executable `output` (found at run
    time through `rpath`, by default this package's lib directory)."""
    tmp = None
    if workdir is None:
        tmp = tempfile.TemporaryDirectory()
        workdir = tmp.name
    try:
        converted = []
        for s in sources:
            base = os.path.splitext(os.path.basename(s))[0]
            dst = os.path.join(workdir, base + ".gpumpi.cpp")
            if s.lower().endswith(".c"):
                converter.convert(s, dst, include_dirs)
            else:
                dst = s  # already device source
            converted.append(dst)
        lib = os.path.join(PKG, "lib")
        cmd = [HIPCC] + FLAGS + (["-fgpu-rdc"] if len(converted) > 1 else [])
        cmd += ["-I", os.path.join(ROOT, "include"), "-I", os.path.join(PKG, "csrc")]
        for d in include_dirs:
            cmd += ["-I", d]
        # the converted copies live elsewhere: quoted includes relative to a
        # source's own directory are found through it
        for d in dict.fromkeys(os.path.dirname(os.path.abspath(x)) for x in sources):
            cmd += ["-I", d]
        cmd += ["-x", "hip"] + converted + ["-o", output, "-L", lib, "-lgpumpi", "-Wl,-rpath," + (rpath or lib)]
        cmd += list(extra)
        if verbose:
            print(" ".join(cmd), file=sys.stderr)
        p = subprocess.run(cmd, capture_output=True, text=True)
        msg = "\n".join(l for l in (p.stdout + p.stderr).splitlines() if "packed-fp32-ops" not in l)
        if p.returncode != 0:
            raise RuntimeError(f"gpumpicc: hipcc failed for {sources}:\n{msg}")
        return output
    finally:
        if tmp is not None:
            tmp.cleanup()


def main(argv=None):
    argv = sys.argv[1:] if argv is None else argv
    if not argv or argv[0] in ("-h", "--help"):
        print(__doc__)
        return 0
    sources, incs, args, out = split_args(argv)
    if not sources:
        print("gpumpicc: no source files", file=sys.stderr)
        return 2
    build(sources, out, incs, args, verbose=bool(os.environ.get("GPUMPICC_VERBOSE")))
    return 0


if __name__ == "__main__":
    sys.exit(main())
