#!/bin/bash
# Builds libobme_hip.so for gfx950 in-tree (next to this script's package).
# Packed FP32 instructions are disabled (-packed-fp32-ops).
set -euo pipefail
HERE="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
OUT="${OBME_OUT:-${HERE}/../oceanbiome/libobme_hip.so}"
HIPCC="${HIPCC:-/opt/rocm/bin/hipcc}"
"${HIPCC}" --offload-arch=gfx950 -O3 -std=c++17 -fPIC -shared \
    -Xclang -target-feature -Xclang -packed-fp32-ops \
    -Wall -Wno-unused-function -Wno-unused-variable \
    ${OBME_EXTRA_FLAGS:-} \
    -o "${OUT}.tmp" "${HERE}/obme_api.hip" "${HERE}/obme_pisces.hip"
mv "${OUT}.tmp" "${OUT}"
echo "built ${OUT}"
