#!/bin/bash
# builds tests/cpp/build/driver_replay against the in-tree library
set -e
D=$(cd "$(dirname "$0")" && pwd)
R=$(cd "$D/../.." && pwd)
mkdir -p "$D/build"
g++ -O2 -std=c++17 -Wall -Wextra -I "$R/include" "$D/driver_replay.cpp" -L "$R/pfem2_amd" -lpfem2_amd \
    -Wl,-rpath,'$ORIGIN/../../../pfem2_amd' -o "$D/build/driver_replay"
