#!/usr/bin/env bash
# Build the reference's Newton-Krylov-hookstep solver and Arnoldi routine
# (test oracles).
#
# Test infrastructure only.  Compiles NewtonHook.f90, GMRESm.f90 and
# arnoldi.f from the sources where they lie (REF, default /root/reference)
# with amdflang, links them with oracle/newton_driver.f90 and
# oracle/arnoldi_driver.f90 against MKL's LAPACK (dgelsy, dgesvd, dgebal,
# dgehrd, dhseqr, dhsein, dgebak; /opt/conda/lib/libmkl_rt.so).  Objects go
# to a temporary directory; only the executables land in
# oracle/_ref/newton_ref and oracle/_ref/arnoldi_ref.
set -euo pipefail
REF=${REF:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
OUT=$HERE/_ref
FC=${FC:-amdflang}
[ -f "$REF/NewtonHook.f90" ] || { echo "reference sources not found in $REF" >&2; exit 2; }
mkdir -p "$OUT"
TMP=$(mktemp -d /tmp/opf_newton_build.XXXXXX)
trap 'rm -rf "$TMP"' EXIT
cd "$TMP"
$FC -O2 -c "$REF/NewtonHook.f90" -o NewtonHook.o
$FC -O2 -c "$REF/GMRESm.f90" -o GMRESm.o
$FC -O2 -c "$HERE/newton_driver.f90" -o newton_driver.o
$FC -o "$OUT/newton_ref" newton_driver.o NewtonHook.o GMRESm.o -L/opt/conda/lib -lmkl_rt \
    -Wl,-rpath,/opt/conda/lib
# newton.f90 includes arnoldi.f into free-form source
$FC -O2 -ffree-form -c "$REF/arnoldi.f" -o arnoldi.o
$FC -O2 -c "$HERE/arnoldi_driver.f90" -o arnoldi_driver.o
$FC -o "$OUT/arnoldi_ref" arnoldi_driver.o arnoldi.o -L/opt/conda/lib -lmkl_rt -Wl,-rpath,/opt/conda/lib
echo "built $OUT/newton_ref $OUT/arnoldi_ref" >&2
