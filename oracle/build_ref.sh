#!/bin/bash
# Build oracle/_ref/libnrqcd_ref.so from the reference's own pure Fortran
# modules where they lie under /root/reference/src, plus oracle/ref_bind.f90.
# Only modules without MPI, PETSc or MKL dependencies are compiled (the rest of
# the reference needs libraries this image lacks and is not built).
# Outputs go to oracle/_ref/ only.  Test infrastructure, not product.
set -euo pipefail
HERE="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
REF="${NRQCD_REFERENCE_SRC:-/root/reference/src}"
OUT="$HERE/_ref"
FC="${FC:-amdflang}"
if [ ! -d "$REF" ]; then
    echo "build_ref: $REF not present; skipping" >&2
    exit 0
fi
mkdir -p "$OUT"
cd "$OUT"
FFLAGS="-O2 -fPIC -module-dir $OUT"
for m in precision mathconstants matrixoperations su3 su2 ranlxd_generator; do
    "$FC" $FFLAGS -c "$REF/$m.f90" -o "$OUT/$m.o"
done
"$FC" $FFLAGS -c "$HERE/ref_bind.f90" -o "$OUT/ref_bind.o"
"$FC" -shared -fPIC -o "$OUT/libnrqcd_ref.so" \
    "$OUT/precision.o" "$OUT/mathconstants.o" "$OUT/matrixoperations.o" \
    "$OUT/su3.o" "$OUT/su2.o" "$OUT/ranlxd_generator.o" "$OUT/ref_bind.o"
echo "build_ref: built $OUT/libnrqcd_ref.so"
