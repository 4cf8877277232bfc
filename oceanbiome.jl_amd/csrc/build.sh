#!/bin/bash
# Builds libobme_hip.so for gfx950 in-tree (next to this script's package):
# each translation unit is compiled to an object in parallel, then linked.
# Packed FP32 instructions are disabled (-packed-fp32-ops).
set -euo pipefail
HERE="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
OUT="${OBME_OUT:-${HERE}/../oceanbiome/libobme_hip.so}"
HIPCC="${HIPCC:-/opt/rocm/bin/hipcc}"
FLAGS=(--offload-arch=gfx950 -O3 -std=c++17 -fPIC
       -Xclang -target-feature -Xclang -packed-fp32-ops
       -Wall -Wno-unused-function -Wno-unused-variable)
TMP="$(mktemp -d)"
trap 'rm -rf "${TMP}"' EXIT
# obme_pisces.hip: FP32 division and square root need not be correctly
# rounded (its FP32 arithmetic is the FP32-mixed path; FP64 is unaffected).
declare -A TU_FLAGS=([obme_api]="" [obme_pisces]="-fno-hip-fp32-correctly-rounded-divide-sqrt")
pids=()
for src in obme_api obme_pisces; do
    "${HIPCC}" "${FLAGS[@]}" ${TU_FLAGS[$src]} ${OBME_EXTRA_FLAGS:-} -c -o "${TMP}/${src}.o" "${HERE}/${src}.hip" &
    pids+=($!)
done
for p in "${pids[@]}"; do wait "$p"; done
"${HIPCC}" --offload-arch=gfx950 -fPIC -shared -o "${OUT}.tmp" "${TMP}/obme_api.o" "${TMP}/obme_pisces.o"
mv "${OUT}.tmp" "${OUT}"
echo "built ${OUT}"
