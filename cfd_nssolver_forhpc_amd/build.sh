#!/bin/sh
# Builds libcfd_hip.so for gfx950 in-tree (cfd_nssolver_forhpc_amd/_lib/).
#   sh build.sh            the library
#   sh build.sh sanitize   libcfd_hip_san.so: the same library with its host
#                          code built with AddressSanitizer and
#                          UndefinedBehaviorSanitizer (device code unchanged;
#                          loaded with clang's ASan runtime preloaded, for the
#                          CPU tests of the host-side entry points)
# Each translation unit compiles in its own hipcc process (in parallel, at
# most BUILD_JOBS at once), then one link step makes the shared library.
set -e
HERE="$(cd "$(dirname "$0")" && pwd)"
mkdir -p "$HERE/_lib" "$HERE/_lib/obj"
HIPCC="${HIPCC:-/opt/rocm/bin/hipcc}"
UNITS="capi k_flow k_ell k_krylov k_sor adjoint k_mg io"
COMMON="--offload-arch=gfx950 -std=c++17 -fPIC -Xclang -target-feature -Xclang -packed-fp32-ops -I $HERE/../include"
if [ "$1" = "sanitize" ]; then
  OUT="$HERE/_lib/libcfd_hip_san.so"
  TAG=san
  FLAGS="$COMMON -O1 -g -fno-omit-frame-pointer -Xarch_host -fsanitize=address -Xarch_host -fsanitize=undefined"
  LINK="-Xarch_host -fsanitize=address -Xarch_host -fsanitize=undefined -Xarch_host -shared-libasan"
else
  OUT="$HERE/_lib/libcfd_hip.so"
  TAG=opt
  FLAGS="$COMMON -O3 -Wall -Wno-unused-function -Wno-unused-variable"
  LINK=""
fi
JOBS="${BUILD_JOBS:-8}"
pids=""
n=0
for u in $UNITS; do
  # shellcheck disable=SC2086
  "$HIPCC" $FLAGS -c "$HERE/csrc/$u.hip" -o "$HERE/_lib/obj/$u.$TAG.o" 2>"$HERE/_lib/obj/$u.$TAG.err" &
  pids="$pids $!"
  n=$((n + 1))
  if [ "$n" -ge "$JOBS" ]; then
    for p in $pids; do wait "$p" || { cat "$HERE"/_lib/obj/*."$TAG".err >&2; exit 1; }; done
    pids=""
    n=0
  fi
done
for p in $pids; do wait "$p" || { cat "$HERE"/_lib/obj/*."$TAG".err >&2; exit 1; }; done
grep -h -v "packed-fp32-ops" "$HERE"/_lib/obj/*."$TAG".err >&2 || true
OBJS=""
for u in $UNITS; do OBJS="$OBJS $HERE/_lib/obj/$u.$TAG.o"; done
# shellcheck disable=SC2086
"$HIPCC" --offload-arch=gfx950 -shared -fPIC $LINK -o "$OUT.tmp" $OBJS
mv "$OUT.tmp" "$OUT"
