#!/bin/sh
# Build the C++ API test programs (host code only; they link the engine
# library and find it through a relative rpath, so the tree can move).
set -e
here=$(cd "$(dirname "$0")" && pwd)
root=$(cd "$here/../.." && pwd)
mkdir -p "$here/build"
for src in lower_line track_ring read_lattice rhic_diagnostics strong_strong; do
    g++ -std=c++17 -O2 -Wall -Wextra -I"$root/include" "$here/$src.cpp" -o "$here/build/$src.tmp" \
        -L"$root/simtrack_amd/lib" -lsimtrack_amd -Wl,-rpath,'$ORIGIN/../../../simtrack_amd/lib'
    mv "$here/build/$src.tmp" "$here/build/$src"
done
