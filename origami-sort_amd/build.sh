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
for src in csrc/sort_kernels.hip csrc/ogs_api.cpp csrc/ogs_host.cpp; do
    obj="_build/$(basename "${src%.*}").o"
    $HIPCC $FLAGS "$@" -c "$src" -o "$obj" &
    pids+=($!)
done
for p in "${pids[@]}"; do wait "$p"; done
$HIPCC --offload-arch=gfx950 -shared -o libogs.so.tmp _build/*.o -lpthread
mv -f libogs.so.tmp libogs.so
