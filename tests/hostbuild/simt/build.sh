#!/bin/bash
# Builds the CPU simulation of the runtime + launcher + rank programs with a
# scenario driver: build.sh OUT DRIVER.cpp [extra flags]
set -e
HERE=$(cd "$(dirname "$0")" && pwd)
ROOT=$(cd "$HERE/../../.." && pwd)
OUT=$1; DRIVER=$2; shift 2
CXX=${CXX:-/opt/rocm/llvm/bin/clang++}
FLAGS="-std=c++17 -g -O1 -fno-omit-frame-pointer -ffp-contract=off -fsanitize=address,undefined -fno-sanitize=alignment,function -mllvm -asan-globals=0 -fsanitize-address-use-after-return=never -I $HERE -I $ROOT/include -I $ROOT/dphpc-gpu_mpi_amd/csrc $*"
mkdir -p "$(dirname "$OUT")/obj"
OBJ="$(dirname "$OUT")/obj"
srcs="$ROOT/dphpc-gpu_mpi_amd/csrc/host/launcher.cpp $ROOT/dphpc-gpu_mpi_amd/csrc/progs/registry.cpp $HERE/simt.cpp $DRIVER"
for p in $ROOT/dphpc-gpu_mpi_amd/csrc/progs/*.hip; do srcs="$srcs $p"; done
objs=""
pids=""
for s in $srcs; do
  o="$OBJ/$(basename $s).o"
  objs="$objs $o"
  cov="-fsanitize-coverage=inline-8bit-counters,trace-loads,trace-stores"
  case "$s" in *simt.cpp) cov="";; esac
  $CXX $FLAGS $cov -x c++ -c "$s" -o "$o" &
  pids="$pids $!"
done
for p in $pids; do wait $p; done
$CXX $FLAGS $objs -o "$OUT"
