#!/bin/bash
# Host AddressSanitizer + UndefinedBehaviorSanitizer build of the library sources and the batch driver
# into one executable (device code is not instrumented: -Xarch_host).
# Run on a GPU machine as:
#   ASAN_OPTIONS=detect_leaks=0 UBSAN_OPTIONS=halt_on_error=1:print_stacktrace=1 tests/host_asan/asan_batch
set -e
HERE=$(cd "$(dirname "$0")" && pwd)
ROOT=$(cd "$HERE/../.." && pwd)
S=$ROOT/rust-sloth_amd/csrc
/opt/rocm/bin/hipcc --offload-arch=gfx950 -O1 -g -std=c++17 -ffp-contract=off -fno-fast-math \
  -Xclang -target-feature -Xclang -packed-fp32-ops \
  -Xarch_host -fsanitize=address -Xarch_host -fsanitize=undefined -Xarch_host -fno-omit-frame-pointer -I "$ROOT/include" -I "$S" \
  "$S/sloth_capi.hip" "$S/sloth_global.hip" "$S/sloth_tiled.hip" "$S/sloth_shade.hip" "$S/sloth_encode.hip" \
  "$S/sloth_models.hip" "$HERE/asan_batch.cpp" -o "$HERE/asan_batch"
