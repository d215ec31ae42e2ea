#!/bin/bash
# Builds tests/hostemu/libhostemu.so: the cell/verify kernel code compiled for the host.
set -e
cd "$(dirname "$0")"
hipcc --offload-arch=gfx950 -O2 -std=c++17 -fPIC -shared -fopenmp \
  -Xclang -target-feature -Xclang -packed-fp32-ops -I../../include \
  cot_hostemu.cpp -o libhostemu.so.tmp
mv libhostemu.so.tmp libhostemu.so
