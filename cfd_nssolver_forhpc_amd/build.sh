#!/bin/sh
# Builds libcfd_hip.so for gfx950 in-tree (cfd_nssolver_forhpc_amd/_lib/).
#   sh build.sh            the library
#   sh build.sh sanitize   libcfd_hip_san.so: the same library with its host
#                          code built with AddressSanitizer and
#                          UndefinedBehaviorSanitizer (device code unchanged;
#                          loaded with clang's ASan runtime preloaded, for the
#                          CPU tests of the host-side entry points)
set -e
HERE="$(cd "$(dirname "$0")" && pwd)"
mkdir -p "$HERE/_lib"
HIPCC="${HIPCC:-/opt/rocm/bin/hipcc}"
SRC="$HERE/csrc/capi.hip $HERE/csrc/k_flow.hip $HERE/csrc/k_ell.hip $HERE/csrc/k_krylov.hip $HERE/csrc/k_sor.hip $HERE/csrc/adjoint.hip $HERE/csrc/k_mg.hip $HERE/csrc/io.hip"
if [ "$1" = "sanitize" ]; then
  OUT="$HERE/_lib/libcfd_hip_san.so"
  # shellcheck disable=SC2086
  "$HIPCC" --offload-arch=gfx950 -O1 -g -std=c++17 -fPIC -shared -fno-omit-frame-pointer \
    -Xclang -target-feature -Xclang -packed-fp32-ops \
    -Xarch_host -fsanitize=address -Xarch_host -fsanitize=undefined -Xarch_host -shared-libasan \
    -I "$HERE/../include" -o "$OUT.tmp" $SRC
else
  OUT="$HERE/_lib/libcfd_hip.so"
  # shellcheck disable=SC2086
  "$HIPCC" --offload-arch=gfx950 -O3 -std=c++17 -fPIC -shared \
    -Xclang -target-feature -Xclang -packed-fp32-ops \
    -Wall -Wno-unused-function -Wno-unused-variable \
    -I "$HERE/../include" \
    -o "$OUT.tmp" $SRC
fi
mv "$OUT.tmp" "$OUT"
