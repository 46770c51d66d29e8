#!/bin/sh
# Builds the FP64 CPU oracle (test infrastructure only) into oracle/libfv3_oracle.so.
# Contraction is off so every expression rounds as written, like the HIP FP64 form.
# OpenMP runs the independent levels (transport) and rows (remap) in parallel;
# results do not depend on the thread count.
set -e
here=$(cd "$(dirname "$0")" && pwd)
${CC:-gcc} -O2 -fPIC -shared -ffp-contract=off -fno-fast-math -std=c11 -fopenmp -Wall -Wno-unused-function \
  -o "$here/libfv3_oracle.so.tmp" "$here/fv3_oracle_tp.c" "$here/fv3_oracle_mapz.c" "$here/fv3_oracle_l2e.c" "$here/fv3_oracle_dyn.c" -lm
mv "$here/libfv3_oracle.so.tmp" "$here/libfv3_oracle.so"
# The upstream fv_tp_2d / fv_mapz sources are absent from the reference
# (atmos_cubed_sphere is an empty submodule), so there is no oracle/_ref build.
