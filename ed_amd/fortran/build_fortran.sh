#!/bin/bash
# build_fortran.sh -- compile the ISO_C_BINDING layer, the GPU-backed QS3
# modules and the two driver programs with ROCm flang, linked against
# ed_amd/libqs3x.so (rpath $ORIGIN/..).  64-bit default integers, so sectors
# with C(NOS, N_down) > 2^31 (L = 1000) are representable.
set -euo pipefail
HERE="$(cd "$(dirname "$0")" && pwd)"
FC=${FC:-/opt/rocm/lib/llvm/bin/flang}
OUT="$HERE/_build"
mkdir -p "$OUT"
FFLAGS="-O2 -fPIC -fdefault-integer-8 -module-dir $OUT -I $OUT"
$FC $FFLAGS -c "$HERE/qs3x_iso_c.f90" -o "$OUT/qs3x_iso_c.o"
$FC $FFLAGS -c "$HERE/qs3x_modules.f90" -o "$OUT/qs3x_modules.o"
$FC $FFLAGS -c "$HERE/qs3x_main.f90" -o "$OUT/qs3x_main.o"
$FC $FFLAGS -c "$HERE/qs3x_dsflan.f90" -o "$OUT/qs3x_dsflan.o"
LIBDIR="$(cd "$HERE/.." && pwd)"
$FC -o "$HERE/qs3x_main.exe" "$OUT/qs3x_main.o" "$OUT/qs3x_modules.o" "$OUT/qs3x_iso_c.o" \
  -L"$LIBDIR" -lqs3x -Wl,-rpath,'$ORIGIN/..'
$FC -o "$HERE/qs3x_dsflan.exe" "$OUT/qs3x_dsflan.o" "$OUT/qs3x_iso_c.o" \
  -L"$LIBDIR" -lqs3x -Wl,-rpath,'$ORIGIN/..'
echo "built $HERE/qs3x_main.exe $HERE/qs3x_dsflan.exe"
