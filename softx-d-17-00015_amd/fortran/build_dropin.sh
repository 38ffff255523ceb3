#!/usr/bin/env bash
# Build the reference's own main program (main.f90, unmodified) against the
# drop-in modules of this directory: the proof that a host written against
# the reference's modules links unchanged.  The reference's parameters.f90
# is the run's configuration file; a configured copy (sizes and the run
# settings given as KEY=VALUE, e.g. i_maxtstep=100 d_timestep=0.01d0) and a
# copy of parallel.h without _CPUTIME are written to a temporary directory
# outside the repository, preprocessed with cpp -traditional-cpp as the
# reference's build does, and compiled with the drop-in modules; only the
# executable lands in fortran/_dropin/ (git-ignored).
#
# usage: fortran/build_dropin.sh NAME N K M [KEY=VALUE ...]
set -euo pipefail
NAME=$1; N=$2; K=$3; M=$4; shift 4
REF=${REF:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
OUT=$HERE/_dropin
FC=${FC:-amdflang}
[ -f "$REF/main.f90" ] || { echo "reference sources not found in $REF" >&2; exit 2; }
mkdir -p "$OUT"
TMP=$(mktemp -d /tmp/opf_dropin_build.XXXXXX)
trap 'rm -rf "$TMP"' EXIT
mkdir -p "$TMP/program" "$TMP/pp" "$TMP/build"
sed -e '/^#define _CPUTIME/d' "$REF/parallel.h" > "$TMP/parallel.h"
SED=(-e "s/\(i_N *= *\)[0-9][0-9]*/\1$N/" -e "s/\(i_K *= *\)[0-9][0-9]*/\1$K/" -e "s/\(i_M *= *\)[0-9][0-9]*/\1$M/")
for kv in "$@"; do
  key=${kv%%=*}; val=${kv#*=}
  SED+=(-e "s/\(\b$key *= *\)[-+0-9.dDeE]*[0-9.]/\1$val/")
done
sed "${SED[@]}" "$REF/parameters.f90" > "$TMP/program/parameters.f90"
grep -E "i_(N|K|M) +=|$(printf '%s|' "$@" | sed -e 's/=[^|]*//g' -e 's/|$//')" "$TMP/program/parameters.f90" >&2 || true
cd "$TMP/program"
cpp -traditional-cpp -P -I"$TMP/program" parameters.f90 > "$TMP/pp/parameters.f90"
cpp -traditional-cpp -P -I"$TMP/program" "$REF/main.f90" > "$TMP/pp/main.f90"
make -s -C "$HERE" PARAMS="$TMP/pp/parameters.f90" BUILD="$TMP/build" FC="$FC"
cd "$TMP/pp"
$FC -O2 -I"$TMP/build" -c main.f90 -o main.o
$FC -O2 -o "$OUT/opf_dropin_$NAME" main.o "$TMP/build/libopf_dropin.a" \
    -L"$HERE/../opf" -lopf -Wl,-rpath,'$ORIGIN/../../opf'
echo "built $OUT/opf_dropin_$NAME (N=$N K=$K M=$M $*)" >&2
