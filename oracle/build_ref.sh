#!/bin/bash
# build_ref.sh -- compile the reference's own hot-path modules into
# oracle/_ref/libqs3ref.so (test infrastructure only; never shipped).
#
# Sources are read where they lie under the reference root (default
# /root/reference, or $QS3_REFERENCE_ROOT).  Nothing is copied into the repo;
# objects, .mod files and the library go to oracle/_ref/ (git-ignored).
#
# Compiled unmodified: source/input_param.f90, my_l1_blas.f90,
# state_lists.f90, ham2vec_v3.f90, get_expectation_values.f90, plus
# oracle/ref_harness.f90 (our own C-callable harness).
#
# -fdefault-integer-8: the shipped default-integer build overflows C(L,N) at
# L=1000 (C(1000,4) > 2^31-1), so the oracle is 64-bit throughout.
# -O2 rather than the reference's -Ofast (configure.ac:22): strict IEEE
# semantics, so the k != 0 norms are comparable bit for bit.
#
# get_expectation_values.f90 also holds calcu_DSF_wavevector, which calls the
# BLAS routines zdotc and dznrm2.  This image has no BLAS under those names and
# the harness never calls calcu_DSF_wavevector, so the two undefined
# references are marked weak (llvm-objcopy --weaken-symbol) instead of being
# linked; nothing stands in for them.  The lanczos and eigen_solver modules need
# LAPACK (dstevx, zheevr, zgemm) and are not built: their parity is pinned by
# closed forms and dense diagonalisation in tests/.
set -euo pipefail
HERE="$(cd "$(dirname "$0")" && pwd)"
REF="${QS3_REFERENCE_ROOT:-/root/reference}"
SRC="$REF/source"
OUT="$HERE/_ref"
FC=/opt/rocm/lib/llvm/bin/flang
OBJCOPY=/opt/rocm/lib/llvm/bin/llvm-objcopy
if [ ! -d "$SRC" ]; then
  echo "build_ref.sh: reference sources not found at $SRC; skipping" >&2
  exit 0
fi
mkdir -p "$OUT"
FFLAGS="-O2 -fopenmp -fPIC -fdefault-integer-8 -module-dir $OUT -I $OUT"
for f in input_param my_l1_blas state_lists ham2vec_v3; do
  $FC $FFLAGS -c "$SRC/$f.f90" -o "$OUT/$f.o"
done
# get_expectation_values.f90 is compiled without -fopenmp: flang 22 lowers its
# "!$omp atomic" update of a complex(8) element (calcu_Sq_a, lines 318 and 340)
# so that the imaginary part of the sum is lost, even with one thread.  Without
# OpenMP the directive is inactive and the loop runs serially, which is the
# reference's intended result.
$FC ${FFLAGS/-fopenmp /} -c "$SRC/get_expectation_values.f90" -o "$OUT/get_expectation_values.o"
$OBJCOPY --weaken-symbol=zdotc_ --weaken-symbol=dznrm2_ "$OUT/get_expectation_values.o"
$FC $FFLAGS -c "$HERE/ref_harness.f90" -o "$OUT/ref_harness.o"
$FC -shared -fopenmp -o "$OUT/libqs3ref.so" \
  "$OUT/input_param.o" "$OUT/my_l1_blas.o" "$OUT/state_lists.o" "$OUT/ham2vec_v3.o" \
  "$OUT/get_expectation_values.o" "$OUT/ref_harness.o" -latomic
echo "built $OUT/libqs3ref.so"

# The U(1)-only modules (source_only_u1/) reuse the module names, so they get
# their own module directory and library.
SRC1="$REF/source_only_u1"
OUT1="$OUT/u1"
if [ -d "$SRC1" ]; then
  mkdir -p "$OUT1"
  FFLAGS1="-O2 -fopenmp -fPIC -fdefault-integer-8 -module-dir $OUT1 -I $OUT1"
  for f in input_param my_l1_blas state_lists ham2vec_v3; do
    $FC $FFLAGS1 -c "$SRC1/$f.f90" -o "$OUT1/$f.o"
  done
  $FC $FFLAGS1 -c "$HERE/ref_harness_u1.f90" -o "$OUT1/ref_harness_u1.o"
  $FC -shared -fopenmp -o "$OUT/libqs3ref_u1.so" \
    "$OUT1/input_param.o" "$OUT1/my_l1_blas.o" "$OUT1/state_lists.o" "$OUT1/ham2vec_v3.o" \
    "$OUT1/ref_harness_u1.o"
  echo "built $OUT/libqs3ref_u1.so"
fi

# DSFLan (source_DSFLan/main.f90): plain Fortran, no BLAS.  Built unmodified as
# an executable with the reference's default integers; it reads
# alpha_beta_values_E0-tmp.d from its working directory and appends
# DSF_itrnum*_eta004.dat there (tests/test_dsflan_reference.py runs it in a
# scratch directory).
SRCD="$REF/source_DSFLan/main.f90"
if [ -f "$SRCD" ]; then
  mkdir -p "$OUT/dsflan"
  $FC -O2 -module-dir "$OUT/dsflan" "$SRCD" -o "$OUT/dsflan.exe"
  echo "built $OUT/dsflan.exe"
fi

# mk_input_list (source_mk_input_list/mk_input_list.f90): the reference's
# square / triangular bond-list generator; reads its namelist from stdin.
SRCM="$REF/source_mk_input_list/mk_input_list.f90"
if [ -f "$SRCM" ]; then
  mkdir -p "$OUT/mkinput"
  $FC -O2 -module-dir "$OUT/mkinput" "$SRCM" -o "$OUT/mk_input_list.exe"
  echo "built $OUT/mk_input_list.exe"
fi

# The reference's own driver (source/main.f90), unmodified, compiled against
# this build's drop-in modules (ed_amd/fortran: input_param, state_lists,
# ham2vec, lanczos, eigen_solver, get_expectation_values over libqs3x.so)
# instead of its CPU modules: the drop-in check that
# tests/test_fortran_driver.py runs on the GPU against the engine's own driver.
SRCMAIN="$SRC/main.f90"
SHIMS="$HERE/../ed_amd/fortran/_build"
if [ -f "$SRCMAIN" ] && [ -f "$SHIMS/qs3x_modules.o" ]; then
  mkdir -p "$OUT/main"
  $FC -O2 -fdefault-integer-8 -module-dir "$OUT/main" -I "$SHIMS" -c "$SRCMAIN" -o "$OUT/main/main.o"
  $FC -o "$OUT/qs3_main_on_qs3x.exe" "$OUT/main/main.o" "$SHIMS/qs3x_modules.o" "$SHIMS/qs3x_iso_c.o" \
    -L"$HERE/../ed_amd" -lqs3x -Wl,-rpath,'$ORIGIN/../../ed_amd'
  echo "built $OUT/qs3_main_on_qs3x.exe"
fi
