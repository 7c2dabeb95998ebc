#!/bin/bash
# Builds the reference Fortran path into oracle/_ref/ (test infrastructure).
#
# Sources are read where they lie (REF, default /root/reference); nothing
# from them is copied into the repository.  The compile-time limits of the
# reference's ewald.h are raised for the larger decks (morbit, mnk, mgrid,
# m_props_in_stack, mstack); that patched header and the copy of
# reptation.f it has to sit beside live only in a temporary directory that
# is removed on exit.  Outputs (binaries only) go to oracle/_ref/:
#   refeval   fixed-configuration / move-replay / timing harness (refeval.f)
#   refqmc    the reference's own driver entry (refmain.f)
# Compiler: amdflang (ROCm LLVM Flang), or gfortran when present.
set -euo pipefail
REF=${REF:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
OUT="$HERE/_ref"
OPT=${OPT:--O3}
if [ ! -f "$REF/reptation.f" ]; then
  echo "build_ref: reference sources not found at $REF; keeping existing oracle/_ref" >&2
  exit 0
fi
FC=${FC:-}
if [ -z "$FC" ]; then
  if command -v gfortran >/dev/null 2>&1; then FC=gfortran
  elif command -v amdflang >/dev/null 2>&1; then FC=amdflang
  elif [ -x /opt/rocm/bin/amdflang ]; then FC=/opt/rocm/bin/amdflang
  else echo "build_ref: no Fortran compiler" >&2; exit 1; fi
fi
TMP=$(mktemp -d /tmp/degas_ref.XXXXXX)
trap 'rm -rf "$TMP"' EXIT
cp "$REF/reptation.f" "$REF/fake_mpi/mpif.f" "$REF/fake_mpi/mpif.h" "$TMP/"
# raised limits: N<=166, 64 orbitals/spin, 2*nk+1<=2048 g-vectors,
# ngrid+nk_ewald+1<=2100 table rows, n_props<=4000 stack slots
sed -e 's/mgrid= 901 ,/mgrid=2100 ,/' \
    -e 's/morbit=  42,mnk=196/morbit=  64,mnk=2048/' \
    -e 's/mstack=10000,/mstack=1000,/' \
    -e 's/m_props_in_stack=900,/m_props_in_stack=4000,/' \
    "$REF/ewald.h" > "$TMP/ewald.h"
for pat in 'mgrid=2100' 'morbit=  64,mnk=2048' 'mstack=1000' 'm_props_in_stack=4000'; do
  grep -q "$pat" "$TMP/ewald.h" || { echo "build_ref: limit patch '$pat' did not apply" >&2; exit 1; }
done
# Fortran does not short-circuit .and.: with der_nskip = 0 (no 'der' deck
# line) amdflang evaluates mod(istp,0) in vmc and rmc (reptation.f:1001,
# 1343) and traps.  Guard the divisor; the branch stays false as intended.
sed -i -e 's/mod(istp,der_nskip)/mod(istp,max(der_nskip,1))/' \
       -e 's/mod(conta,der_nskip)/mod(conta,max(der_nskip,1))/' "$TMP/reptation.f"
[ "$(grep -c 'max(der_nskip,1)' "$TMP/reptation.f")" = 2 ] || { echo "build_ref: der_nskip guard did not apply" >&2; exit 1; }
cp "$HERE/refeval.f" "$HERE/refmain.f" "$TMP/"
cd "$TMP"
FFLAGS="$OPT -w"
$FC $FFLAGS -c reptation.f -o reptation.o 2> build.log || { cat build.log >&2; exit 1; }
$FC $FFLAGS -c mpif.f -o mpif.o
$FC $FFLAGS -c refeval.f -o refeval.o
$FC $FFLAGS -c refmain.f -o refmain.o
mkdir -p "$OUT"
$FC $FFLAGS refeval.o reptation.o mpif.o -o "$OUT/refeval"
$FC $FFLAGS refmain.o reptation.o mpif.o -o "$OUT/refqmc"
# drop-in build: the reference driver with its compute_properties replaced
# by the GPU forwarder (2degas_amd/fortran); needs lib2degas_hip.so
PKG="$(dirname "$HERE")/2degas_amd"
if [ -f "$PKG/lib2degas_hip.so" ]; then
  sed -e 's/^      subroutine compute_properties(ifdist)/      subroutine compute_properties_cpu(ifdist)/' \
      reptation.f > reptation_dropin.f
  grep -q 'subroutine compute_properties_cpu' reptation_dropin.f || { echo "build_ref: drop-in rename failed" >&2; exit 1; }
  $FC $FFLAGS -c reptation_dropin.f -o reptation_dropin.o
  $FC $FFLAGS -c "$PKG/fortran/qmc_hip_mod.f90" -o qmc_hip_mod.o
  cp "$PKG/fortran/compute_properties_hip.f" .
  $FC $FFLAGS -c compute_properties_hip.f -o compute_properties_hip.o
  $FC $FFLAGS refmain.o reptation_dropin.o mpif.o compute_properties_hip.o qmc_hip_mod.o \
      -L"$PKG" -l2degas_hip -Wl,-rpath,'$ORIGIN/../../2degas_amd' -Wl,-rpath,/opt/rocm/lib \
      -o "$OUT/refqmc_hip"
  echo "build_ref: built $OUT/refqmc_hip (reference driver + GPU compute_properties)"
  # and with the reference's vmc replaced by the device VMC driver
  # (2degas_amd/fortran/vmc_hip.f, the Fortran caller of qmc_vmc_sweep)
  sed -e 's/^      subroutine vmc$/      subroutine vmc_cpu/' reptation_dropin.f > reptation_dropvmc.f
  grep -q '^      subroutine vmc_cpu$' reptation_dropvmc.f || { echo "build_ref: vmc rename failed" >&2; exit 1; }
  $FC $FFLAGS -c reptation_dropvmc.f -o reptation_dropvmc.o
  cp "$PKG/fortran/vmc_hip.f" .
  $FC $FFLAGS -c vmc_hip.f -o vmc_hip.o
  $FC $FFLAGS refmain.o reptation_dropvmc.o mpif.o compute_properties_hip.o vmc_hip.o qmc_hip_mod.o \
      -L"$PKG" -l2degas_hip -L/opt/rocm/lib -lamdhip64 -Wl,-rpath,'$ORIGIN/../../2degas_amd' \
      -Wl,-rpath,/opt/rocm/lib -o "$OUT/refqmc_hipvmc"
  echo "build_ref: built $OUT/refqmc_hipvmc (reference driver + GPU compute_properties + GPU vmc)"
fi
echo "$FC $($FC --version 2>/dev/null | head -1) $OPT" > "$OUT/COMPILER"
echo "build_ref: built $OUT/refeval $OUT/refqmc with $FC $OPT"
