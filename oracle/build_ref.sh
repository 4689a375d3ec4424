#!/usr/bin/env bash
# Builds the reference's own solver modules, unchanged and in place, into
# oracle/_ref/ref_driver. Test infrastructure only: the driver is the checker
# that pins the C restatement (oracle/sph_oracle.c) to the reference.
#
# Every reference source is compiled where it lies under REF_ROOT. Left out:
# src/Initialize/getValueFromToml.f90 (needs toml-f, not in this image) and
# src/main.f90 (replaced by oracle/ref_driver.f90, which sets the globals the
# toml reader would set). No reference file is copied.
#
# Flags follow the reference build (CMakeLists.txt:14-21, :37-52): -cpp,
# OpenMP; -O2 on the default x86-64 target, so no FMA contraction is possible.
set -euo pipefail
HERE="$(cd "$(dirname "$0")" && pwd)"
REF_ROOT="${REF_ROOT:-/root/reference}"
OUT="$HERE/_ref"
FC="${FC:-/opt/rocm/bin/amdflang}"
if [ ! -d "$REF_ROOT/src" ]; then
    echo "build_ref: $REF_ROOT/src not found; skipping reference build" >&2
    exit 0
fi
mkdir -p "$OUT/obj"
S="$REF_ROOT/src"
# dependency order of the Fortran modules
SRCS="tools.f90 Initialize/controlDict.f90 Initialize/SPH.f90 NNPS/graphics.f90
Numeric/kernel.f90 NNPS/asymmetricSearch.f90 NNPS/backGroundGrid.f90 NNPS/nnps.f90
Physic/eos.f90 Physic/viscosity.f90 Physic/density.f90 Physic/internalForce.f90
Physic/externalForce.f90 Numeric/artificialViscosity.f90 Numeric/artificialHeat.f90
Numeric/artificialTerm.f90 Numeric/divergence.f90 Numeric/kernelGradientCorrection.f90
Numeric/shifting.f90 Numeric/hsml.f90 Numeric/correctVelocity.f90
Numeric/courantNumber.f90 Numeric/area.f90 Numeric/decomposePar.f90 Physic/HE.f90
Physic/dummyParticles.f90 Output/output.f90 TimeIntegration/timeIntegration.f90
Physic/boundaryCondition.f90 TimeIntegration/singleStep.f90 Initialize/input.f90"
FFLAGS="-cpp -fopenmp -O2 -w -I$S -module-dir $OUT/obj"
OBJS=""
for f in $SRCS; do
    o="$OUT/obj/$(basename "$f" .f90).o"
    "$FC" $FFLAGS -c "$S/$f" -o "$o"
    OBJS="$OBJS $o"
done
"$FC" -cpp -fopenmp -O2 -w -I"$OUT/obj" -module-dir "$OUT/obj" -c "$HERE/ref_driver.f90" -o "$OUT/obj/ref_driver.o"
"$FC" -fopenmp -O2 $OBJS "$OUT/obj/ref_driver.o" -o "$OUT/ref_driver"
echo "build_ref: built $OUT/ref_driver"

# Drop-in link: the same reference program with single_step taken from
# mysph_amd/fortran/single_step_gpu.f90 instead of singleStep.f90, so the
# reference's own leap_frog drives the GPU derivatives through libsphgpu.
REPO="$(cd "$HERE/.." && pwd)"
if [ -f "$REPO/mysph_amd/libsphgpu.so" ]; then
    F="$REPO/mysph_amd/fortran"
    "$FC" -cpp -O2 -module-dir "$OUT/obj" -c "$F/sphgpu_m.f90" -o "$OUT/obj/sphgpu_m.o"
    "$FC" -cpp -fopenmp -O2 -w -I"$OUT/obj" -module-dir "$OUT/obj" -c "$F/single_step_gpu.f90" \
        -o "$OUT/obj/single_step_gpu.o"
    GOBJS=""
    for o in $OBJS; do
        [ "$(basename "$o")" = "singleStep.o" ] || GOBJS="$GOBJS $o"
    done
    "$FC" -fopenmp -O2 $GOBJS "$OUT/obj/ref_driver.o" "$OUT/obj/sphgpu_m.o" "$OUT/obj/single_step_gpu.o" \
        -L"$REPO/mysph_amd" -lsphgpu -Wl,-rpath,'$ORIGIN/../../mysph_amd' -o "$OUT/ref_driver_gpu"
    echo "build_ref: built $OUT/ref_driver_gpu"
fi
