#!/bin/bash
# Host AddressSanitizer + UndefinedBehaviorSanitizer build of libpyqed_amd's host code with the driver
# tests/sanitize/host_driver.cpp:  bash tests/sanitize/build.sh OUT_DIR
# Device code is built without sanitizers (each -fsanitize= on the hipcc lines applies to the host pass
# only); the driver and the link are host-only clang++ lines that name no HIP source.
set -euo pipefail
here="$(cd "$(dirname "$0")" && pwd)"
root="$(cd "$here/../.." && pwd)"
out="${1:?output directory}"
mkdir -p "$out"
HIPCC="${HIPCC:-/opt/rocm/bin/hipcc}"
SAN=(-Xarch_host -fsanitize=address -Xarch_host -fsanitize=undefined -Xarch_host -fno-omit-frame-pointer
     -Xarch_host -fno-sanitize-recover=undefined)
for src in heom.hip capi.cpp; do
  "$HIPCC" -O1 -g -std=c++17 -fPIC --offload-arch=gfx950 -I "$root/include" -I "$root/pyqed_amd/csrc" \
    -Xclang -target-feature -Xclang -packed-fp32-ops "${SAN[@]}" -x hip -c "$root/pyqed_amd/csrc/$src" \
    -o "$out/$src.o"
done
# the driver and the link use the same LLVM as hipcc, whose sanitizer runtimes match the instrumentation
CXX="${SAN_CXX:-/opt/rocm/llvm/bin/clang++}"
"$CXX" -O1 -g -std=c++17 -fsanitize=address -fsanitize=undefined -fno-omit-frame-pointer -I "$root/include" \
  -I /opt/rocm/include -D__HIP_PLATFORM_AMD__ -c "$here/host_driver.cpp" -o "$out/host_driver.o"
"$CXX" -fsanitize=address -fsanitize=undefined "$out/host_driver.o" "$out/heom.hip.o" "$out/capi.cpp.o" \
  -L /opt/rocm/lib -Wl,-rpath,/opt/rocm/lib -lamdhip64 -o "$out/host_driver"
