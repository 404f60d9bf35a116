#!/bin/bash
# Builds libogs.so (gfx950) in-tree.  Usage: build.sh [extra hipcc flags]
set -euo pipefail
HERE="$(cd "$(dirname "$0")" && pwd)"
cd "$HERE"
HIPCC=${HIPCC:-/opt/rocm/bin/hipcc}
FLAGS="--offload-arch=gfx950 -O3 -std=c++17 -fPIC -Wall -Wno-unused-function -Xclang -target-feature -Xclang -packed-fp32-ops"
mkdir -p _build
rm -f _build/*.o
pids=()
for src in csrc/sort_kernels.hip csrc/kway_kernels.hip csrc/gen_kernels.hip csrc/ogs_noalloc.hip csrc/ogs_api.cpp csrc/ogs_host.cpp csrc/ogs_serve.cpp; do
    obj="_build/$(basename "${src%.*}").o"
    $HIPCC $FLAGS "$@" -c "$src" -o "$obj" &
    pids+=($!)
done
for p in "${pids[@]}"; do wait "$p"; done
$HIPCC --offload-arch=gfx950 -shared -o libogs.so.tmp _build/*.o -lpthread -ldl
mv -f libogs.so.tmp libogs.so
# drop-in check program for include/origami_gpu_compat.h (run by tests/test_gpu_compat.py)
g++ -O2 -std=c++17 -I ../include ../tests/compat_bench_sort.cpp -o compat_bench_sort.tmp \
    -L . -logs -ldl -Wl,-rpath,'$ORIGIN' -L /opt/rocm/lib -lamdhip64
mv -f compat_bench_sort.tmp compat_bench_sort
