#!/bin/bash
# TEST INFRASTRUCTURE ONLY: builds the reference solver as the oracle.
#
# Compiles the reference FORTRAN77 solver from its own two source files under
# /root/reference with amdflang, linked with ref_harness.f (a replacement main
# program that dumps the COMMON blocks) and ref_dump.c.  The only change to the
# reference is the one its header documents (multall-open-18.3.f:77-80): the
# array dimensions in the PARAMETER statement of commall-open-18.3, which must
# shrink to fit the 2 GiB static-data limit of the x86-64 small code model.
# The two sources are copied into a scratch directory under oracle/_ref/ only
# for the compile (FORTRAN INCLUDE resolves next to the including file first)
# and deleted afterwards; only the executables stay in oracle/_ref/.
#
# Usage: oracle/build_ref.sh [fp32|fp64|fp32w|fp32l|fp64l|fp32c|fp64c|all]
# (all = fp32, fp64, fp32w; the long builds fp32l/fp64l, JM <= 1300 with
# IM, KM <= 38, are made on request for the longest test decks, and
# fp32c/fp64c, IM <= 97, KM <= 81, for the centrifugal-stage workload)
set -euo pipefail
HERE="$(cd "$(dirname "$0")" && pwd)"
SRC=/root/reference/Multall-open-18.3/multall-18.3
OUT="$HERE/_ref"
FC=${FC:-amdflang}
DIMS="ID=65,JD=600,KD=65,MAXKI=65,NRS=21,IG1=30,"
# a second FP32 build for pitchwise-wide decks (IM <= 96, JM <= 200, KM <= 33)
DIMS_WIDE="ID=96,JD=200,KD=33,MAXKI=96,NRS=21,IG1=30,"
# long builds for decks of many rows (IM, KM <= 38, JM <= 1300)
DIMS_LONG="ID=38,JD=1300,KD=38,MAXKI=38,NRS=21,IG1=30,"
# builds for the centrifugal-stage workload (IM <= 97, JM <= 290, KM <= 81;
# 32 level-1 multigrid blocks pitchwise at IM = 97)
DIMS_CENT="ID=97,JD=290,KD=81,MAXKI=97,NRS=21,IG1=40,"
if [ ! -f "$SRC/multall-open-18.3.f" ]; then
  echo "build_ref: reference sources not present; skipping" >&2
  exit 0
fi
mkdir -p "$OUT"

build_one() {
  local v=$1 flags=$2 dims=${3:-$DIMS}
  local B="$OUT/build_$v"
  rm -rf "$B"; mkdir -p "$B"
  trap "rm -rf $B" ERR
  cp "$SRC/multall-open-18.3.f" "$SRC/commall-open-18.3" "$B/"
  sed -i "s/PARAMETER(ID=55,JD=2500,KD=74,MAXKI=74,NRS=21,IG1=30,/PARAMETER($dims/" "$B/commall-open-18.3"
  grep -q "PARAMETER($dims" "$B/commall-open-18.3"
  cp "$HERE/ref_harness.f" "$B/"
  (cd "$B" && $FC -O3 $flags -c multall-open-18.3.f -o mf.o \
     && $FC -O3 $flags -c ref_harness.f -o harness.o)
  gcc -O2 -c "$HERE/ref_dump.c" -o "$B/dump.o"
  # the reference's own main program is replaced by the harness's
  objcopy --weaken-symbol=main --weaken-symbol=_QQmain "$B/mf.o"
  $FC -O3 $flags "$B/harness.o" "$B/mf.o" "$B/dump.o" -o "$OUT/multall_ref_$v"
  rm -rf "$B"
  echo "built $OUT/multall_ref_$v"
}

case "${1:-all}" in
  fp32) build_one fp32 "" ;;
  fp64) build_one fp64 "-fdefault-real-8 -fdefault-double-8" ;;
  fp32w) build_one fp32w "" "$DIMS_WIDE" ;;
  fp32l) build_one fp32l "" "$DIMS_LONG" ;;
  fp64l) build_one fp64l "-fdefault-real-8 -fdefault-double-8" "$DIMS_LONG" ;;
  fp32c) build_one fp32c "" "$DIMS_CENT" ;;
  fp64c) build_one fp64c "-fdefault-real-8 -fdefault-double-8" "$DIMS_CENT" ;;
  all)  build_one fp32 "" & p1=$!
        build_one fp64 "-fdefault-real-8 -fdefault-double-8" & p2=$!
        build_one fp32w "" "$DIMS_WIDE" & p3=$!
        wait $p1; wait $p2; wait $p3 ;;
esac
