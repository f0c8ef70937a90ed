#!/usr/bin/env bash
# Builds oracle/_ref/ref_harness from the reference's CPU-path sources where
# they lie (the CPU_RUN subset of SU(N)/Makefile:70-79: random.cpp and
# suncpp/*.cpp), plus oracle/ref_harness.cpp.  Flags follow the reference
# Makefile (-g -O2, Makefile:30,37) with -D CPU_RUN (Makefile:41-43).
# The OpenCL headers the reference's algebra headers include
# (clinterface/clinterface.h:40-41) come from the system include path; no
# OpenCL library is linked because the CPU_RUN path calls none.
#
# Test infrastructure only: outputs go to oracle/_ref/ (git-ignored) and are
# used to write the golden fixtures under tests/golden/.  Skips silently (exit
# 0 with a message) when the reference tree is absent.
set -euo pipefail
HERE="$(cd "$(dirname "$0")" && pwd)"
REF="${QCD_REFERENCE_ROOT:-/root/reference}/SU(N)"
OUT="$HERE/_ref"
if [ ! -d "$REF" ]; then
    echo "build_ref: reference tree not present, skipping"
    exit 0
fi
mkdir -p "$OUT"
SRCS=(
    "$REF/random/random.cpp"
    "$REF/suncpp/su2/algebra_su2.cpp"
    "$REF/suncpp/su2/update_su2.cpp"
    "$REF/suncpp/su3/algebra_su3.cpp"
    "$REF/suncpp/su3/update_su3.cpp"
    "$REF/suncpp/coord_work/coord_work.cpp"
)
g++ -D CPU_RUN -D CL_TARGET_OPENCL_VERSION=120 -g -O2 -w \
    -I "$REF" "${SRCS[@]}" "$HERE/ref_harness.cpp" -o "$OUT/ref_harness"
echo "build_ref: built $OUT/ref_harness"
# Second CPU-baseline row: the same sources at -O3 with AVX2/FMA code
# (x86-64-v3).  -march=native would target the CPU of the machine that
# builds it, which need not be the one that runs the benchmark; x86-64-v3
# runs on every AVX2 host.
g++ -D CPU_RUN -D CL_TARGET_OPENCL_VERSION=120 -O3 -march=x86-64-v3 -w \
    -I "$REF" "${SRCS[@]}" "$HERE/ref_harness.cpp" -o "$OUT/ref_harness_o3"
echo "build_ref: built $OUT/ref_harness_o3"
