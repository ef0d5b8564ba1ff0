#!/bin/bash
# Compiles the reference's Fortran kernels, in place under /root/reference,
# into oracle/_ref/ (one shared library per f2py module, as the reference's
# build.py does) with the ROCm LLVM Fortran compiler.  Used only to write and
# check golden vectors (tests/golden/make_golden.py); the sources are read
# where they lie and nothing is copied.  No reference build system is run.
#
# Flags: -O2 -ffp-contract=off so that the arithmetic follows the source
# text without fused multiply-adds (the reference's own -march=native build
# may contract them, which these checks do not follow).  fortran_diag.f90 and
# gmg/fortran_multigrid.f90 are fixed-form sources.
set -euo pipefail
REF=${REF:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
OUT="$HERE/_ref"
FC=${FC:-/opt/rocm/bin/amdflang}
if [ ! -d "$REF/core" ]; then
  echo "build_ref.sh: $REF not present, skipping" >&2
  exit 0
fi
if ! command -v "$FC" >/dev/null 2>&1; then
  echo "build_ref.sh: $FC not found, skipping" >&2
  exit 0
fi
mkdir -p "$OUT"
FFLAGS="-O2 -ffp-contract=off -fPIC"
build() { # name source [extra flags]
  local name=$1 src=$2; shift 2
  "$FC" $FFLAGS "$@" -c "$src" -o "$OUT/$name.o" 2>"$OUT/$name.log"
  "$FC" -shared -o "$OUT/lib$name.so" "$OUT/$name.o"
  rm -f "$OUT/$name.o"
}
build fortran_advection "$REF/core/fortran_advection.f90"
build fortran_fluxes "$REF/core/fortran_fluxes.f90"
build fortran_operators "$REF/core/fortran_operators.f90"
build fortran_diag "$REF/core/fortran_diag.f90" -ffixed-form
build fortran_multigrid "$REF/core/gmg/fortran_multigrid.f90" -ffixed-form
echo "build_ref.sh: built $(ls "$OUT"/*.so | wc -l) libraries in $OUT"
