#!/bin/sh
# Builds libcfd_hip.so for gfx950 in-tree (cfd_nssolver_forhpc_amd/_lib/).
set -e
HERE="$(cd "$(dirname "$0")" && pwd)"
mkdir -p "$HERE/_lib"
HIPCC="${HIPCC:-/opt/rocm/bin/hipcc}"
"$HIPCC" --offload-arch=gfx950 -O3 -std=c++17 -fPIC -shared \
  -Xclang -target-feature -Xclang -packed-fp32-ops \
  -Wall -Wno-unused-function -Wno-unused-variable \
  -I "$HERE/../include" \
  -o "$HERE/_lib/libcfd_hip.so.tmp" \
  "$HERE"/csrc/capi.hip "$HERE"/csrc/k_flow.hip "$HERE"/csrc/k_ell.hip "$HERE"/csrc/k_krylov.hip "$HERE"/csrc/k_sor.hip "$HERE"/csrc/adjoint.hip "$HERE"/csrc/k_mg.hip "$HERE"/csrc/io.hip
mv "$HERE/_lib/libcfd_hip.so.tmp" "$HERE/_lib/libcfd_hip.so"
