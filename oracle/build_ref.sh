#!/bin/bash
# Builds the reference pysic Fortran core, straight from the read-only
# reference sources, plus this directory's ref_driver.f90, into
# oracle/_ref/libpysic_ref.so.  Checker only (see ref_driver.f90).
# Needs /root/reference; does nothing (exit 0) where it is absent.
set -euo pipefail
HERE="$(cd "$(dirname "$0")" && pwd)"
SRC=/root/reference/fortran
OUT="$HERE/_ref"
if [ ! -d "$SRC" ]; then
  echo "build_ref: $SRC absent, skipping the reference build"
  exit 0
fi
FC=${FC:-/opt/rocm/lib/llvm/bin/amdflang}
mkdir -p "$OUT/obj"
cd "$OUT/obj"
# serial build (no -DMPI): MPI.f90 supplies serial stubs, is_my_atom == true
FFLAGS="-cpp -O2 -fPIC"
for f in Mersenne.F90 MPI.f90 Quaternions.f90 Utility.f90 Geometry.f90 Potentials.f90 Core.f90 PyInterface.f90; do
  "$FC" $FFLAGS -c "$SRC/$f"
done
"$FC" $FFLAGS -c "$HERE/ref_driver.f90"
"$FC" -shared -fPIC -o "$OUT/libpysic_ref.so.tmp" ./*.o
mv "$OUT/libpysic_ref.so.tmp" "$OUT/libpysic_ref.so"
echo "build_ref: wrote $OUT/libpysic_ref.so"

# MPI build (-DMPI: MPI.f90 includes mpif.h and distributes the atoms over
# the ranks, Core.f90 sums the forces with mpi_allreduce), against the MPICH
# of the image when it has one.  Used only as the multi-core CPU baseline.
MPI_ROOT=${MPI_ROOT:-/opt/conda}
if [ -f "$MPI_ROOT/include/mpif.h" ] && [ -f "$MPI_ROOT/lib/libmpifort.so" ]; then
  mkdir -p "$OUT/obj_mpi"
  cd "$OUT/obj_mpi"
  MFLAGS="-cpp -DMPI -I$MPI_ROOT/include -O2 -fPIC"
  for f in Mersenne.F90 MPI.f90 Quaternions.f90 Utility.f90 Geometry.f90 Potentials.f90 Core.f90 PyInterface.f90; do
    "$FC" $MFLAGS -c "$SRC/$f"
  done
  "$FC" $MFLAGS -c "$HERE/ref_driver.f90"
  "$FC" -shared -fPIC -o "$OUT/libpysic_ref_mpi.so.tmp" ./*.o -L"$MPI_ROOT/lib" \
    -Wl,-rpath,"$MPI_ROOT/lib" -lmpifort -lmpi
  mv "$OUT/libpysic_ref_mpi.so.tmp" "$OUT/libpysic_ref_mpi.so"
  echo "build_ref: wrote $OUT/libpysic_ref_mpi.so"
else
  echo "build_ref: no MPI under $MPI_ROOT, skipping the MPI build"
fi
