#!/bin/bash
# Host build of the library for AddressSanitizer checks: the HIP sources are
# compiled as C++ against tests/cpuemu/hip/hip_runtime.h, which runs every GPU
# thread of a workgroup as a host thread (workgroups one after another).
# -fsanitize=bounds also checks every index into a fixed-size array member,
# which ASan cannot: kernel arguments are structs passed by value, and their
# copies sit in stack memory without redzones.
#   tests/cpuemu/build.sh OUT.so
set -e
D=$(cd "$(dirname "$0")"; pwd)
R=$(cd "$D/../.."; pwd)
C=$R/specfempp_amd/csrc
OUT=$1
T=$(mktemp -d)
FLAGS="-std=c++20 -g -O1 -fPIC -fsanitize=address,bounds -fno-sanitize-recover=bounds -fno-omit-frame-pointer -I$D -I$C -I$R/include"
for f in engine batch block3d kernels acoustic general3d; do
  g++ $FLAGS -x c++ -c "$C/$f.hip" -o "$T/$f.o" &
done
g++ $FLAGS -c "$C/assembly.cpp" -o "$T/assembly.o" &
wait
g++ -shared -fsanitize=address,bounds -o "$OUT" "$T"/engine.o "$T"/batch.o "$T"/kernels.o "$T"/acoustic.o "$T"/general3d.o "$T"/block3d.o "$T"/assembly.o -lpthread
rm -rf "$T"
