#!/usr/bin/env bash
# Build the reference Openpipeflow core (test oracle) for one resolution.
#
# Test infrastructure only.  Compiles the reference's own core modules from
# the sources where they lie (REF, default /root/reference) with amdflang and
# links them with oracle/ref_driver.f90 against MKL's LAPACK and FFTW3
# wrappers (/opt/conda/lib/libmkl_rt.so).  The reference keeps its sizes as
# compile-time parameters (parameters.f90:17-20), so a configured copy of
# parameters.f90 with i_N/i_K/i_M/i_Mp replaced is written to a temporary
# directory outside the repository; the copy, the preprocessed sources and the
# objects are deleted afterwards.  Only the executable lands in oracle/_ref/.
#
# usage: oracle/build_ref.sh NAME N K M [Mp] [SYMM]
#   SYMM: one of mirrorsym, shiftrefl, shiftrott sets that b_* flag of
#   parameters.f90:26-28 (vel_imposesym then projects every step)
# With NR and/or NS in the environment (NR*NS > 1) the core is built for
# that many MPI ranks (_Nr, _Ns of parallel.h:5-6, a configured copy) against
# MPICH's mpif.h and libraries in /opt/conda, driven by
# oracle/ref_mpi_driver.f90; the executable is oracle/_ref/opf_ref_mpi_NAME.
# With PLUGIN=nonnewtonian or PLUGIN=les the plugin's module and procedures
# are cut from the reference's plugin source by line range (the module,
# init_params_model, and get_nu .. update_nonlin_terms; its util, file
# output and the main.f90 it includes are left out) and the driver's
# var_null(2) points call them through oracle/plugin_PLUGIN.f90.
set -euo pipefail
NAME=$1; N=$2; K=$3; M=$4; MP=${5:-1}; SYMM=${6:-}
NR=${NR:-1}; NS=${NS:-1}; PLUGIN=${PLUGIN:-none}
REF=${REF:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
OUT=$HERE/_ref
FC=${FC:-amdflang}
[ -f "$REF/velocity.f90" ] || { echo "reference sources not found in $REF" >&2; exit 2; }
mkdir -p "$OUT"
TMP=$(mktemp -d /tmp/opf_ref_build.XXXXXX)
trap 'rm -rf "$TMP"' EXIT
mkdir -p "$TMP/program" "$TMP/pp"
# The sources include "../parallel.h" relative to their own directory.
sed -e '/^#define _CPUTIME/d' -e "s/^#define _Nr .*/#define _Nr $NR/" -e "s/^#define _Ns .*/#define _Ns $NS/" \
    "$REF/parallel.h" > "$TMP/parallel.h"
sed -e "s/\(i_N *= *\)[0-9][0-9]*/\1$N/" \
    -e "s/\(i_K *= *\)[0-9][0-9]*/\1$K/" \
    -e "s/\(i_M *= *\)[0-9][0-9]*/\1$M/" \
    -e "s/\(i_Mp *= *\)[0-9][0-9]*/\1$MP/" \
    "$REF/parameters.f90" > "$TMP/program/parameters.f90"
if [ -n "$SYMM" ]; then
  sed -i -e "s/\(b_$SYMM *= *\)\.false\./\1.true./" "$TMP/program/parameters.f90"
  grep -E "b_$SYMM +=" "$TMP/program/parameters.f90" >&2
fi
grep -E "i_(N|K|M|Mp) +=" "$TMP/program/parameters.f90" | head -4 >&2
cd "$TMP/program"
for f in mpi meshs variables transform.fftw3 timestep velocity; do
  cpp -traditional-cpp -P -I"$TMP/program" "$REF/$f.f90" > "$TMP/pp/$f.f90"
done
cpp -traditional-cpp -P -I"$TMP/program" parameters.f90 > "$TMP/pp/parameters.f90"
PLUGOBJ=""
# the plugin sources get parallel.h's loop macros through the main.f90 they
# include; the cut-out procedures include it directly
echo '#include "../parallel.h"' > "$TMP/program/plugin.f90"
case "$PLUGIN" in
  none) ;;
  nonnewtonian)   # module 8-27, init_params_model 62-86, get_nu + update_nonlin_terms 92-247
    sed -n -e '8,27p' -e '62,86p' -e '92,247p' "$REF/nonnewtonian.f90" >> "$TMP/program/plugin.f90" ;;
  les)            # module 20-34, init_params_model 69-96, get_nu .. update_nonlin_terms 102-339
    sed -n -e '20,34p' -e '69,96p' -e '102,339p' "$REF/LES.f90" >> "$TMP/program/plugin.f90" ;;
  *) echo "unknown PLUGIN $PLUGIN" >&2; exit 2 ;;
esac
if [ "$PLUGIN" != none ]; then
  cpp -traditional-cpp -P -I"$TMP/program" "$TMP/program/plugin.f90" > "$TMP/pp/plugin.f90"
  PLUGOBJ="plugin.o"
fi
cd "$TMP/pp"
FFLAGS="-O2"
# the module arrays of the largest sizes exceed 2 GiB of static data
if [ $((N * K * M)) -gt 1000000 ]; then FFLAGS="$FFLAGS -mcmodel=medium"; fi
DRIVER=ref_driver; EXE="$OUT/opf_ref_$NAME"; LIBS="-lmkl_rt"
if [ $((NR * NS)) -gt 1 ]; then
  FFLAGS="$FFLAGS -I/opt/conda/include"
  DRIVER=ref_mpi_driver; EXE="$OUT/opf_ref_mpi_$NAME"; LIBS="-lmkl_rt -lmpifort -lmpi"
fi
$FC $FFLAGS -c "$REF/cheby.f" -o cheby.o
for f in parameters mpi meshs variables transform.fftw3 timestep velocity; do
  $FC $FFLAGS -c $f.f90 2>/dev/null
done
if [ -n "$PLUGOBJ" ]; then $FC $FFLAGS -c plugin.f90 -o plugin.o; fi
$FC $FFLAGS -c "$HERE/plugin_$PLUGIN.f90" -o hook.o
$FC $FFLAGS -c "$HERE/$DRIVER.f90" -o driver.o
$FC $FFLAGS -o "$EXE" driver.o hook.o $PLUGOBJ parameters.o mpi.o cheby.o meshs.o \
    variables.o transform.fftw3.o timestep.o velocity.o \
    -L/opt/conda/lib $LIBS -Wl,-rpath,/opt/conda/lib
echo "built $EXE (N=$N K=$K M=$M Mp=$MP ranks=${NR}x${NS})" >&2
