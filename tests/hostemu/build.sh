#!/bin/bash
# Builds tests/hostemu/libhostemu.so: the cell/verify kernel code compiled for the host.
# With SAN=1, builds the same code with AddressSanitizer and UndefinedBehaviorSanitizer on the
# host side only, into the path given as the first argument (load it with clang's ASan runtime
# preloaded).
set -e
cd "$(dirname "$0")"
if [ "${SAN:-0}" = "1" ]; then
  out="$1"
  hipcc --offload-arch=gfx950 -O1 -g -std=c++17 -fPIC -shared -fopenmp \
    -Xclang -target-feature -Xclang -packed-fp32-ops -I../../include \
    -Xarch_host -fsanitize=address -Xarch_host -fsanitize=undefined \
    -Xarch_host -fno-sanitize-recover=undefined -Xarch_host -fno-omit-frame-pointer \
    cot_hostemu.cpp -o "$out"
  exit 0
fi
hipcc --offload-arch=gfx950 -O2 -std=c++17 -fPIC -shared -fopenmp \
  -Xclang -target-feature -Xclang -packed-fp32-ops -I../../include \
  cot_hostemu.cpp -o libhostemu.so.tmp
mv libhostemu.so.tmp libhostemu.so
