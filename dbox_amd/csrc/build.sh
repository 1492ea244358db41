#!/bin/bash
# Builds libdbox_hip.so for gfx950 in place (dbox_amd/libdbox_hip.so).
# -ffp-contract=off keeps the reference's float rounding (no FMA contraction);
# packed fp32 instructions are disabled for every HIP translation unit; loop
# heads start on 64-byte lines (the solver's level loops run a few percent
# faster than wherever the surrounding code happens to place them).
set -euo pipefail
cd "$(dirname "$0")"
HIPCC=${HIPCC:-/opt/rocm/bin/hipcc}
ARCH=${DBX_ARCH:-gfx950}
FLAGS="--offload-arch=$ARCH -O3 -std=c++17 -ffp-contract=off -fPIC -Wall -Wno-unused-function -Xclang -target-feature -Xclang -packed-fp32-ops -falign-loops=64"
OUT=../libdbox_hip.so
OBJ=_build
mkdir -p $OBJ
pids=()
for src in dbx_kernels.hip dbx_runtime.hip dbx_mlp.hip; do
  $HIPCC $FLAGS -c $src -o $OBJ/${src%.hip}.o 2> >(grep -v "packed-fp32-ops" >&2) &
  pids+=($!)
done
for p in "${pids[@]}"; do wait $p; done
$HIPCC --offload-arch=$ARCH -shared -fPIC -o $OUT.tmp $OBJ/dbx_kernels.o $OBJ/dbx_runtime.o $OBJ/dbx_mlp.o
mv $OUT.tmp $OUT
echo "built $(realpath $OUT)"
