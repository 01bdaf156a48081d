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
# Incremental: an object is rebuilt when it is missing, older than its
# source or than any runtime, include or simulation header, or when the
# flags changed; an executable is relinked when older than an object.
stamp="$OUT/.flags"
if [ ! -f "$stamp" ] || [ "$(cat "$stamp")" != "$CXX $FLAGS $COV" ]; then
  rm -f "$OBJ"/*.o
  echo "$CXX $FLAGS $COV" > "$stamp"
fi
newest_hdr=$(find "$ROOT/include" "$ROOT/dphpc-gpu_mpi_amd/csrc" "$HERE/hip" "$HERE/build.sh" \
        \( -name '*.h' -o -name '*.hpp' -o -name 'build.sh' \) -printf '%T@ %p\n' | sort -n | tail -1 | cut -d' ' -f2-)
stale() {  # stale OBJECT SOURCE
  [ ! -f "$1" ] && return 0
  [ "$2" -nt "$1" ] && return 0
  [ "$newest_hdr" -nt "$1" ] && return 0
  return 1
}
pids=""
objs=""
# A driver may be several sources joined by ':' (a program of several
# translation units); the executable is named after the first.
for s in $common $(echo "$@" | tr ':' ' '); do
  o="$OBJ/$(basename $s).o"
  if stale "$o" "$s"; then
    $CXX $FLAGS $COV -x c++ -c "$s" -o "$o" &
    pids="$pids $!"
  fi
done
# the scheduler itself is not instrumented (its own accesses are not lanes')
if stale "$OBJ/simt.cpp.o" "$HERE/simt.cpp"; then
  $CXX $FLAGS -c "$HERE/simt.cpp" -o "$OBJ/simt.cpp.o" &
  pids="$pids $!"
fi
for p in $pids; do wait $p; done
for s in $common; do objs="$objs $OBJ/$(basename $s).o"; done
for d in "$@"; do
  dobjs=""
  for part in $(echo "$d" | tr ':' ' '); do dobjs="$dobjs $OBJ/$(basename $part).o"; done
  first=$(echo "$d" | cut -d: -f1)
  exe="$OUT/$(basename $first .cpp)"
  relink=0
  [ -f "$exe" ] || relink=1
  for ob in $objs "$OBJ/simt.cpp.o" $dobjs; do [ "$ob" -nt "$exe" ] && relink=1; done
  [ $relink = 1 ] && $CXX $FLAGS $objs "$OBJ/simt.cpp.o" $dobjs -o "$exe"
done
exit 0
