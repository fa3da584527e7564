#!/usr/bin/env bash
# Build libbfgpu.so (gfx950) in-tree from csrc/ (one hipcc per source file,
# run in parallel).
set -euo pipefail
here="$(cd "$(dirname "$0")" && pwd)"
mkdir -p "$here/lib"
HIPCC="${HIPCC:-/opt/rocm/bin/hipcc}"
FLAGS=(--offload-arch=gfx950 -O3 -std=c++17 -fPIC -Wall -Wno-unused-result
       -Xclang -target-feature -Xclang -packed-fp32-ops)
objs=()
pids=()
for src in "$here"/csrc/*.hip; do
  obj="$here/lib/$(basename "${src%.hip}").o"
  "$HIPCC" "${FLAGS[@]}" -c "$src" -o "$obj" &
  pids+=("$!")
  objs+=("$obj")
done
status=0
for p in "${pids[@]}"; do
  wait "$p" || status=1
done
if [ "$status" -ne 0 ]; then
  rm -f "${objs[@]}"
  echo "build.sh: compilation failed" >&2
  exit 1
fi
"$HIPCC" --offload-arch=gfx950 -shared -fPIC -o "$here/lib/libbfgpu.so.tmp" "${objs[@]}"
mv "$here/lib/libbfgpu.so.tmp" "$here/lib/libbfgpu.so"
rm -f "${objs[@]}"
