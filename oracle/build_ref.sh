#!/bin/bash
# Build the reference SORD solver (serial process module + OpenMP) from the
# Fortran sources where they lie under /root/reference/sord, with ROCm's
# amdflang.  Outputs go only to oracle/_ref/{r4,r8}/sord.x (FP32 and FP64,
# mirroring sord/Makefile.in:16-18 REALSIZE=8 -> -fdefault-real-8).
#
# One token of sord/diff_cn_op.f90:10 is patched on the fly (the file is piped
# through sed into the compiler, never copied): `real :: h, b` declares `b`
# both real and integer (:10-11), which every Fortran compiler rejects.
#
# Compile order follows the `use` dependencies (arrays needs process, so the
# list in cst/sord.py:188-216 is not a valid order on its own).
set -euo pipefail
REF=${SORD_REF:-/root/reference/sord}
HERE=$(cd "$(dirname "$0")" && pwd)
FC=${FC:-/opt/rocm/lib/llvm/bin/amdflang}
if [ ! -d "$REF" ]; then
    echo "reference sources not present ($REF); skipping reference build" >&2
    exit 0
fi
ORDER="globals kernels diff_cn_op diff_nc_op hourglass_op boundary_cond surf_normals \
utilities fortran_io thread_omp process_serial arrays input_output statistics setup \
grid_generation material_model boundary_pml kinematic_source material_resample \
time_integration stress dynamic_rupture acceleration sord"
for rs in 4 8; do
    out="$HERE/_ref/r$rs"
    mkdir -p "$out"
    flags="-O3 -fopenmp"
    if [ "$rs" = 8 ]; then flags="$flags -fdefault-real-8"; fi
    objs=""
    for m in $ORDER; do
        src="$REF/$m.f90"
        if [ "$m" = diff_cn_op ]; then
            sed '10s/^real :: h, b$/real :: h/' "$src" | \
                (cd "$out" && $FC $flags -c -x f95 - -o "$m.o")
        else
            (cd "$out" && $FC $flags -c "$src" -o "$m.o")
        fi
        objs="$objs $m.o"
    done
    (cd "$out" && $FC $flags -o sord.x $objs)
    echo "built $out/sord.x"
done
