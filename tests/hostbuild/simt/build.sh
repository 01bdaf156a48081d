#!/bin/bash
# Builds the CPU simulation of the device runtime, the host launcher and rank
# programs, and links one executable per scenario driver:
#   build.sh OUTDIR DRIVER.cpp[:MORE.cpp...] [DRIVER.cpp ...]   -> OUTDIR/<driver name>
# SIMT_PROGS (optional): space-separated program sources (progs/) to build
# instead of every progs/*.hip file.
# SIMT_CFLAGS (optional): extra compiler flags for every source.
set -e
HERE=$(cd "$(dirname "$0")" && pwd)
ROOT=$(cd "$HERE/../../.." && pwd)
OUT=$1; shift
CXX=${CXX:-/opt/rocm/llvm/bin/clang++}
FLAGS="-std=c++17 -g -O1 -fno-omit-frame-pointer -ffp-contract=off -fsanitize=address,undefined -fno-sanitize=alignment,function -mllvm -asan-globals=0 -fsanitize-address-use-after-return=never -I $HERE -I $ROOT/include -I $ROOT/dphpc-gpu_mpi_amd/csrc $SIMT_CFLAGS"
COV="-fsanitize-coverage=inline-8bit-counters,trace-loads,trace-stores"
OBJ="$OUT/obj"
mkdir -p "$OBJ"
common="$ROOT/dphpc-gpu_mpi_amd/csrc/host/launcher.cpp $ROOT/dphpc-gpu_mpi_amd/csrc/host/service.cpp $ROOT/dphpc-gpu_mpi_amd/csrc/progs/registry.cpp"
if [ -n "$SIMT_PROGS" ]; then
  for p in $SIMT_PROGS; do common="$common $ROOT/dphpc-gpu_mpi_amd/csrc/progs/$p"; done
else
  for p in $ROOT/dphpc-gpu_mpi_amd/csrc/progs/*.hip; do common="$common $p"; done
fi
pids=""
objs=""
# A driver may be several sources joined by ':' (a program of several
# translation units); the executable is named after the first.
for s in $common $(echo "$@" | tr ':' ' '); do
  o="$OBJ/$(basename $s).o"
  $CXX $FLAGS $COV -x c++ -c "$s" -o "$o" &
  pids="$pids $!"
done
# the scheduler itself is not instrumented (its own accesses are not lanes')
$CXX $FLAGS -c "$HERE/simt.cpp" -o "$OBJ/simt.cpp.o" &
pids="$pids $!"
for p in $pids; do wait $p; done
for s in $common; do objs="$objs $OBJ/$(basename $s).o"; done
for d in "$@"; do
  dobjs=""
  for part in $(echo "$d" | tr ':' ' '); do dobjs="$dobjs $OBJ/$(basename $part).o"; done
  first=$(echo "$d" | cut -d: -f1)
  $CXX $FLAGS $objs "$OBJ/simt.cpp.o" $dobjs -o "$OUT/$(basename $first .cpp)"
done
