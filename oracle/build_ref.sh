#!/usr/bin/env bash
# Build the reference's own CPU path into oracle/_ref/libdawref.so.
#
# TEST INFRASTRUCTURE ONLY.  Compiles the hot-path sources of the reference
# where they lie (read-only tree, default /root/reference) with AMD flang,
# plus our harness oracle/ref_harness.f90.  Flags follow the reference's
# src/Makefile.inc:22 (-O3 -cpp); -fopenmp is dropped because the sources hold
# no OpenMP directive.  No MPI is needed: the path's files do not use it.  The
# reference's `module parameters` lives in main_redatum_DAW.f90:5-37 next to a
# program that uses MPI, so the module's lines are compiled alone from a
# scratch copy under a temporary directory (never committed, never shipped).
# Only the final shared library is written into oracle/_ref/.
set -euo pipefail
REF=${DAW_REFERENCE_ROOT:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
OUT="$HERE/_ref"
if [ ! -d "$REF/src" ]; then
  echo "build_ref: reference tree $REF/src not present; skipping" >&2
  exit 0
fi
FC=${FC:-flang}
FFLAGS="-O3 -cpp -fPIC -w"
TMP=$(mktemp -d /tmp/dawref.XXXXXX)
trap 'rm -rf "$TMP"' EXIT
mkdir -p "$OUT"
cd "$TMP"
sed -n '/^module parameters/,/^end module parameters/p' "$REF/src/main_redatum_DAW.f90" > parameters.f90
SRCS="parameters.f90 $REF/src/kwsinc.f90 $REF/src/libfourier.f90 $REF/src/minresModule.f90 \
      $REF/src/sublinearminres.f90 $REF/src/submodeling2.f90 $REF/src/subwavefieldresample.f90 \
      $REF/src/subresiduals.f90 $REF/src/subname.f90 $REF/src/subacqui.f90 $REF/src/subresample_hicks.f90 \
      $HERE/ref_harness.f90"
OBJS=""
for f in $SRCS; do
  o=$(basename "$f" .f90).o
  $FC $FFLAGS -c "$f" -o "$o"
  OBJS="$OBJS $o"
done
$FC -shared -fPIC $OBJS -o "$OUT/libdawref.so.tmp"
mv "$OUT/libdawref.so.tmp" "$OUT/libdawref.so"
echo "build_ref: wrote $OUT/libdawref.so"
