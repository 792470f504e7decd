#!/bin/sh
# Build the CPU oracle (test infrastructure; see st_oracle.c).
#   liboracle.so        project libm (st_libm.h): bit-comparable with the GPU
#   liboracle_glibc.so  C library libm: what the reference itself calls
#   liboracle_glibc_{trig,asin,exp,pow}.so  the C library for one function
#                       family only (sin/cos/tan, asin, exp, pow)
#   liboracle_opcount.so  the maps' algorithmic operation counts (opcount.cpp)
#   liboracle_dual.so   the maps' exact Jacobians (dualmap.cpp)
set -e
here=$(cd "$(dirname "$0")" && pwd)
root=$(dirname "$here")
mkdir -p "$here/build"
CFLAGS="-std=gnu99 -O2 -ffp-contract=off -mfma -fopenmp -fPIC -shared -fvisibility=hidden -Wall -Wno-unused-function"
INC="-I$root/include -I$root/simtrack_amd/csrc"
gcc $CFLAGS $INC "$here/st_oracle.c" -o "$here/build/liboracle.so.tmp" -lm
mv "$here/build/liboracle.so.tmp" "$here/build/liboracle.so"
gcc $CFLAGS -DORACLE_GLIBC $INC "$here/st_oracle.c" -o "$here/build/liboracle_glibc.so.tmp" -lm
mv "$here/build/liboracle_glibc.so.tmp" "$here/build/liboracle_glibc.so"
for v in trig:1 asin:2 exp:4 pow:8; do
  name=${v%%:*}; mask=${v##*:}
  gcc $CFLAGS -DORACLE_GLIBC_MASK=$mask $INC "$here/st_oracle.c" -o "$here/build/liboracle_glibc_$name.so.tmp" -lm
  mv "$here/build/liboracle_glibc_$name.so.tmp" "$here/build/liboracle_glibc_$name.so"
done
# op-counting build (C++: st_oracle.c with a tracked scalar type), for the F_turn test
g++ -x c++ -std=c++17 -O1 -ffp-contract=off -mfma -fPIC -shared -fvisibility=hidden -w $INC -I"$here" \
    "$here/opcount.cpp" -o "$here/build/liboracle_opcount.so.tmp"
mv "$here/build/liboracle_opcount.so.tmp" "$here/build/liboracle_opcount.so"
# first-order (dual-number) build, the exact Jacobians of the maps (dualmap.cpp)
g++ -x c++ -std=c++17 -O2 -ffp-contract=off -mfma -fPIC -shared -fvisibility=hidden -w $INC -I"$here" \
    "$here/dualmap.cpp" -o "$here/build/liboracle_dual.so.tmp"
mv "$here/build/liboracle_dual.so.tmp" "$here/build/liboracle_dual.so"
