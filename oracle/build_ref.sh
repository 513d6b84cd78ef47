#!/usr/bin/env bash
# Builds oracle/_ref/ref_harness: the reference solver's own hot-path sources,
# compiled where they lie under the reference tree, linked with the harness
# driver oracle/ref_harness.cpp.  Output goes to oracle/_ref/ only.
# Release flags as the reference's Release preset (-O3 -DNDEBUG, C++20).
set -euo pipefail
REF=${CFD_REFERENCE:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
OUT="$HERE/_ref"
MPICXX=${MPICXX:-/opt/conda/bin/mpicxx}
if [ ! -d "$REF/src" ]; then
    echo "reference tree not found at $REF; skipping the reference build" >&2
    exit 0
fi
mkdir -p "$OUT/obj"
export MPICH_CXX=${MPICH_CXX:-g++}
FLAGS="-std=c++20 -O3 -DNDEBUG -I$REF/include -I$REF/external/tomlplusplus/include"
SRCS="
src/mpi/log.cpp
src/mpi/mpi_util.cpp
src/mesh/aux_connectivity.cpp
src/mesh/aux_geometry.cpp
src/mesh/geometry.cpp
src/mesh/reorder.cpp
src/mesh/validate.cpp
src/numerics/gradient/gradient_operator.cpp
src/numerics/gradient/gradient_manager.cpp
src/linalg/bicgstab.cpp
src/linalg/preconditioners.cpp
src/linalg/bsr_matrix.cpp
src/linalg/csr_matrix.cpp
src/linalg/ldu_matrix.cpp
src/linalg/vector.cpp
src/linalg/vector_layout.cpp
src/linalg/types.cpp
src/fields/halo.cpp
src/io/vtk/vtu.cpp
"
OBJS=""
pids=()
for s in $SRCS; do
    o="$OUT/obj/$(echo "$s" | tr '/' '_' | sed 's/\.cpp$/.o/')"
    OBJS="$OBJS $o"
    if [ ! -f "$o" ] || [ "$REF/$s" -nt "$o" ]; then
        "$MPICXX" $FLAGS -c "$REF/$s" -o "$o" &
        pids+=($!)
    fi
done
for p in "${pids[@]}"; do wait "$p"; done
"$MPICXX" $FLAGS -c "$HERE/ref_harness.cpp" -o "$OUT/obj/ref_harness.o"
# Link with the system g++ and its own static libstdc++ (the MPI wrapper's
# library path carries an older libstdc++); only the MPI C library is needed.
g++ -o "$OUT/ref_harness" "$OUT/obj/ref_harness.o" $OBJS -static-libstdc++ -static-libgcc \
    /opt/conda/lib/libmpi.so -Wl,-rpath,/opt/conda/lib
echo "built $OUT/ref_harness"
# the reference's configuration parsers (toml++ 3.4.0 is vendored under external/)
CFG_SRCS="src/io/toml/toml_utilities.cpp src/solver/compressible/config.cpp src/bc/config.cpp"
CFG_OBJS=""
for s in $CFG_SRCS; do
    o="$OUT/obj/$(echo "$s" | tr '/' '_' | sed 's/\.cpp$/.o/')"
    CFG_OBJS="$CFG_OBJS $o"
    if [ ! -f "$o" ] || [ "$REF/$s" -nt "$o" ]; then
        "$MPICXX" $FLAGS -c "$REF/$s" -o "$o"
    fi
done
"$MPICXX" $FLAGS -c "$HERE/ref_config.cpp" -o "$OUT/obj/ref_config.o"
g++ -o "$OUT/ref_config" "$OUT/obj/ref_config.o" $CFG_OBJS "$OUT/obj/src_mpi_log.o" -static-libstdc++ -static-libgcc \
    /opt/conda/lib/libmpi.so -Wl,-rpath,/opt/conda/lib
echo "built $OUT/ref_config"
