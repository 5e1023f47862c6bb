#!/bin/bash
# Builds the reference MPI code (3DMHD-CfCA/MCF3DSGincfca3.f90, linked with
# 3DMHD-CfCA/selfG_MPIfull8.f90 which only provides GRAVTY) from the sources
# where they lie under /root/reference, with AMD flang and the MPICH of
# /opt/conda.  Test infrastructure: the binary goes to oracle/_ref/bin/ and is
# used only to write golden fixtures in this container.
#
# usage: oracle/build_ref.sh [NDX [dtf] [chem] [grav]]  (NDX = compile-time block size, default 18;
#        dtf: read the inhomogeneous inflow density field DTF/D{rank}.dat;
#        chem: keep the reference's chemistry-network Fcool;
#        grav: self-gravity runs, the multigrid arrays zero-filled on entry)
#
# -O2 without -march: baseline x86-64 has no FMA, so nothing is fused.
set -euo pipefail
NDX=${1:-18}
shift || true
VARIANT="$*"
SUFFIX=""
for v in $VARIANT; do SUFFIX="${SUFFIX}_$v"; done
REF=/root/reference/3DMHD-CfCA
HERE=$(cd "$(dirname "$0")" && pwd)
OUT=$HERE/_ref
FC=/opt/rocm/bin/amdflang
MPI=/opt/conda
if [ ! -f "$REF/MCF3DSGincfca3.f90" ]; then
  echo "reference sources not present; skipping reference build" >&2
  exit 0
fi
mkdir -p "$OUT/bin"
TMP=$(mktemp -d "$OUT/tmp_build.XXXXXX")
trap 'rm -rf "$TMP"' EXIT
python3 "$HERE/patch_ref.py" "$REF/MCF3DSGincfca3.f90" "$TMP/m3.f90" "$NDX" $VARIANT
case " $VARIANT " in
  *" grav "*) python3 "$HERE/patch_ref.py" --sg8 "$REF/selfG_MPIfull8.f90" "$TMP/sg8.f90" ;;
  *) cp "$REF/selfG_MPIfull8.f90" "$TMP/sg8.f90" ;;
esac
cd "$TMP"
$FC -O2 -I$MPI/include -c m3.f90 -o m3.o 2>build.log || { cat build.log; exit 1; }
$FC -O2 -I$MPI/include -c sg8.f90 -o sg8.o 2>>build.log || { cat build.log; exit 1; }
$FC -O2 -o "$OUT/bin/mcf_ref_ndx$NDX$SUFFIX" m3.o sg8.o -L$MPI/lib -Wl,-rpath,$MPI/lib -lmpifort -lmpi
echo "built $OUT/bin/mcf_ref_ndx$NDX$SUFFIX"
