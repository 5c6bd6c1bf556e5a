#!/bin/bash
# Build the reference Falkner-Skan-Cooke similarity solver (bls/fsc.f, one
# self-contained source file) from /root/reference into oracle/_ref/fsc.
# Test infrastructure only: it writes the fsc.dat base-flow table that the
# reference bla reads (rbla.f) for the boundary-layer fixtures, and pins the
# library's own similarity solver.
#
#   oracle/build_fsc.sh
#
# Flags of bls/Makefile's default target (-fdefault-real-8 -O3), with ROCm
# amdflang in place of gfortran (not installed here).
set -euo pipefail
REF=${SIMSON_REFERENCE:-/root/reference}/bls
HERE=$(cd "$(dirname "$0")" && pwd)
DEST=$HERE/_ref
FC=${FC:-amdflang}
if [ ! -f "$REF/fsc.f" ]; then
    echo "reference sources not found at $REF" >&2
    exit 3
fi
mkdir -p "$DEST"
SCRATCH=$(mktemp -d /tmp/simson_fsc.XXXXXX)
trap 'rm -rf "$SCRATCH"' EXIT
cd "$SCRATCH"
$FC -fdefault-real-8 -O3 -o fsc "$REF/fsc.f"
cp fsc "$DEST/fsc"
echo "built $DEST/fsc"
