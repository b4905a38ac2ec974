#!/usr/bin/env bash
# Build recipe for oracle/_ref: the CUEVA reference compiled from its own
# sources under /root/reference, driven by oracle/refbuild/ref_harness.f95.
#
# Test infrastructure only.  The reference is configured at compile time
# through Fortran parameters (parameters.f95:10-334); this script writes an
# edited copy of parameters.f95 into a scratch directory outside the
# repository, compiles every unit of the reference Makefile (Makefile:1-13)
# except the driver main_cueva_code.f95 straight from /root/reference, links
# them with the harness and leaves only the binary in oracle/_ref/<name>/.
# No reference source is written into the repository.
#
# usage: build_ref.sh NAME IMAX JMAX ORDER FLUX_SOLVER REC_MP BOUND SIGMA CFL LX LY
#   FLUX_SOLVER 1=Lax-Friedrichs, 2=HLL+MCL, 4=HLLC+MCL, 5=HLL 6=HLLE 7=HLLC (+MPx);  REC_MP 0=MP5 1=MP7 2=MP9
# environment: SLOP LIMIT (default 2), PRIMITIVE_REC MIRK REC_PRIM (default 0) set the
# parameters of the same names.
#   ORDER 1 (MIRK=1 or first-order IMEX) 2=SSP2(2,2,2) or MIRK2 3=SSP2(3,3,2)-LUM 4=SSP3(4,3,3);  BOUND 1 copy, 2 periodic, 6 periodic-y/copy-x
set -euo pipefail

if [ $# -ne 11 ]; then
  echo "usage: $0 NAME IMAX JMAX ORDER FLUX_SOLVER REC_MP BOUND SIGMA CFL LX LY" >&2
  exit 2
fi
NAME=$1 IMAX=$2 JMAX=$3 ORDER=$4 FS=$5 RECMP=$6 BOUND=$7 SIGMA=$8 CFL=$9 LX=${10} LY=${11}

REF=${CUEVA_REFERENCE:-/root/reference/cueva_2D_base}
HERE=$(cd "$(dirname "$0")" && pwd)
OUT=$(cd "$HERE/.." && pwd)/_ref/$NAME
FC=${FC:-/opt/rocm/lib/llvm/bin/flang}
FFLAGS=${FFLAGS:--O2 -fopenmp -mcmodel=large}

if [ ! -d "$REF" ]; then
  echo "reference sources not found at $REF" >&2
  exit 3
fi
if [ "$IMAX" -lt "$JMAX" ]; then
  # the MP scratch u(-6:imax+6,1:68,-1:imax+1) (scalar.f95:371) overflows when jmax > imax
  echo "the reference needs imax >= jmax (got $IMAX x $JMAX)" >&2
  exit 4
fi

WORK=$(mktemp -d "${TMPDIR:-/tmp}/cueva_ref_${NAME}.XXXXXX")
trap 'rm -rf "$WORK"' EXIT

# Edited parameters module: only compile-time constants change.
sed -E \
  -e "s/^( *integer,parameter +:: +order +=) *[0-9]+/\1 $ORDER/" \
  -e "s/^( *integer,parameter +:: +BOUND +=) *[0-9]+/\1 $BOUND/" \
  -e "s/^( *integer,parameter +:: +flux_solver +=) *[0-9]+/\1 $FS/" \
  -e "s/^( *integer,parameter +:: +rec_mp +=) *[0-9]+/\1 $RECMP/" \
  -e "s/^( *integer, parameter :: imax +=) *[0-9]+/\1 $IMAX/" \
  -e "s/^( *integer, parameter :: jmax +=) *[0-9]+/\1 $JMAX/" \
  -e "s/^( *doubleprecision,parameter +:: +Lx +=) *[0-9.d]+/\1 $LX/" \
  -e "s/^( *doubleprecision,parameter +:: +Ly +=) *[0-9.d]+/\1 $LY/" \
  -e "s/^( *doubleprecision,parameter +:: +sigma_0 +=) *[0-9.d]+/\1 $SIGMA/" \
  -e "s/^( *doubleprecision,parameter +:: +CFL +=) *[0-9.d]+/\1 $CFL/" \
  -e "s/^( *integer,parameter +:: +primitive_rec +=) *[0-9]+/\1 ${PRIMITIVE_REC:-0}/" \
  -e "s/^( *integer,parameter +:: +slop +=) *[0-9]+/\1 ${SLOP:-2}/" \
  -e "s/^( *integer,parameter +:: +limit +=) *[0-9]+/\1 ${LIMIT:-2}/" \
  -e "s/^( *integer,parameter +:: +MIRK +=) *[0-9]+/\1 ${MIRK:-0}/" \
  -e "s/^( *integer,parameter +:: +REC_PRIM +=) *[0-9]+/\1 ${REC_PRIM:-0}/" \
  "$REF/parameters.f95" > "$WORK/parameters.f95"

# Every substitution must have hit exactly its line.
check() { grep -Eq "$1" "$WORK/parameters.f95" || { echo "parameter edit failed: $1" >&2; exit 5; }; }
check "order +=  *$ORDER( |$)"
check "BOUND +=  *$BOUND( |$)"
check "flux_solver +=  *$FS( |$)"
check "rec_mp +=  *$RECMP( |$)"
check "imax +=  *$IMAX( |$)"
check "jmax +=  *$JMAX( |$)"
check "Lx +=  *$LX( |$)"
check "Ly +=  *$LY( |$)"
check "sigma_0 +=  *$SIGMA( |$)"
check "CFL +=  *$CFL( |$)"
check "primitive_rec +=  *${PRIMITIVE_REC:-0}( |$)"
check "slop +=  *${SLOP:-2}( |$)"
check "limit +=  *${LIMIT:-2}( |$)"
check "MIRK +=  *${MIRK:-0}( |$)"
check "REC_PRIM +=  *${REC_PRIM:-0}( |$)"

cd "$WORK"
UNITS="scalar threevectors fourvectors funciones imex_scheme mp5 mp7 mp9
  01_openfiles 01_open_backup 02_init_interface 03_init_var_primitive 03_restart_primitive
  04_boundary 05_reconstruct 06_conserved_var 07_clean_var 08_swap1 09_sumint 10_varint
  11_elecvarint 12_varprimitive 12_varprimitive_nr 13_intercon_var 14_flowinit 15_lax_flow
  16_hll_flow_mcl 16_hll_flow_mpx 16_hllc_flow_mcl 16_hllc_flow_mpx 16_hlle_flow_mcl
  16_hlle_flow_mpx 17_fluxandsource 18_sumfin 19_varfin 20_electricfield 21_swap2 22_writer
  23_fin_interface 24_boundary_conserved 24_boundary_flux_conserved 25_boundary_electric
  26_source_boundary"

$FC $FFLAGS -c parameters.f95 -o parameters.o
OBJS="parameters.o"
for u in $UNITS; do
  $FC $FFLAGS -c "$REF/$u.f95" -o "$u.o"
  OBJS="$OBJS $u.o"
done
$FC $FFLAGS -c "$HERE/ref_harness.f95" -o ref_harness.o
mkdir -p "$OUT"
$FC $FFLAGS $OBJS ref_harness.o -o "$OUT/cueva_ref"
printf '%s\n' "imax=$IMAX jmax=$JMAX order=$ORDER flux_solver=$FS rec_mp=$RECMP bound=$BOUND sigma=$SIGMA cfl=$CFL lx=$LX ly=$LY" > "$OUT/config.txt"
echo "built $OUT/cueva_ref"
