#!/bin/bash
# Test infrastructure only.  Builds the reference Fortran solver from its own
# sources, where they lie in the reference tree, into oracle/_ref/, and links
# it with oracle/ref_dump.f90 (a driver of ours that replaces the reference's
# `program main` so per-call arrays can be dumped).
#
# Toolchain: ROCm amdflang with the reference's FP64 promotion flags
# (src/Makefile:43), the image's MPICH Fortran bindings (libfmpich, mpif.h
# under /opt/conda) and MKL for the LAPACK calls (dgesvd, la_math.f90).
# Nothing of the reference is copied; objects and the binary stay in
# oracle/_ref/, which git ignores.
set -euo pipefail
REF=${RC_REFERENCE_DIR:-/root/reference}/src
HERE=$(cd "$(dirname "$0")" && pwd)
OUT=$HERE/_ref
CONDA=${RC_CONDA_PREFIX:-/opt/conda}
FC=${FC:-amdflang}
FFLAGS="-O3 -fdefault-real-8 -fdefault-double-8"

if [ ! -d "$REF" ]; then
    echo "reference sources not found at $REF" >&2
    exit 2
fi
mkdir -p "$OUT/obj"
cd "$OUT/obj"
# compile order of src/Makefile:2-42 (modules first)
FILES="module.f90 parallel.f90 combustion_modeling/species.f90 combustion_modeling/fuel.f90
combustion_modeling/oxidizer.f90 combustion_modeling/thermo_property.f90
combustion_modeling/transport_property.f90 combustion_modeling/diffusion_velocity.f90
combustion_modeling/chemical_kinetics.f90 rnn.f90 period.f90 allocate_arrary.f90 flux_lax.f90
flux_viscous.f90 fvm2d.f90 gaussInteg4.f90 ghostcell.f90 gradient.f90 mesh.f90 initialization.f90
initParam.f90 output.f90 reconstruction_initialization.f90 limiter.f90 shockdetect.f90 startup.f90
Time_RK_LUSGS.f90 Time_LUSGS_steady.f90 Time_RungeKutta.f90 TimeStep.f90 sourceterm.f90 residual.f90
monitor.f90 common_boundary.f90 update_flow_variables.f90 la_math.f90 Con_to_Pri_Jacobian.f90
Time_Trapezoidal.f90 Time_BackwardEuler.f90 Time_ImplicitMidpoint.f90 Time_BDF2.f90"
OBJS=""
for f in $FILES; do
    o=$(basename "$f" .f90).o
    $FC $FFLAGS -I"$CONDA/include" -c "$REF/$f" -o "$o" 2>/dev/null
    OBJS="$OBJS $o"
done
# the reference's main program stays in the object but becomes local, so the
# dump driver's own main program links in its place
objcopy --localize-symbol=main --localize-symbol=_QQmain fvm2d.o fvm2d_lib.o
OBJS=${OBJS/ fvm2d.o / fvm2d_lib.o }
$FC $FFLAGS -I"$CONDA/include" -c "$HERE/ref_dump.f90" -o ref_dump.o
$FC -O3 -o "$OUT/rc_ref_dump" ref_dump.o $OBJS -L"$CONDA/lib" -lfmpich -lmpich -lmkl_rt \
    -Wl,-rpath,"$CONDA/lib" -Wl,-z,execstack
# the reference's own program, unchanged, for whole-run comparisons and timing
$FC -O3 -o "$OUT/rc_ref_main" $(echo $OBJS | sed 's/fvm2d_lib.o/fvm2d.o/') -L"$CONDA/lib" -lfmpich -lmpich -lmkl_rt \
    -Wl,-rpath,"$CONDA/lib" -Wl,-z,execstack
echo "built $OUT/rc_ref_dump $OUT/rc_ref_main"
# the reference host driving the MI355X library (needs the built library)
PKG=$(cd "$HERE/.." && pwd)/rocketcombustor_amd
if [ -f "$PKG/librocketcombustor.so" ]; then
    $FC $FFLAGS -c "$HERE/../include/rocketcombustor_iface.f90" -o rocketcombustor_iface.o
    $FC $FFLAGS -I"$CONDA/include" -c "$HERE/ref_host_gpu.f90" -o ref_host_gpu.o
    $FC -O3 -o "$OUT/rc_ref_host_gpu" ref_host_gpu.o rocketcombustor_iface.o $OBJS -L"$CONDA/lib" -lfmpich -lmpich \
        -lmkl_rt -L"$PKG" -lrocketcombustor -Wl,-rpath,"$CONDA/lib" -Wl,-rpath,'$ORIGIN/../../rocketcombustor_amd' \
        -Wl,-z,execstack
    echo "built $OUT/rc_ref_host_gpu"
fi
