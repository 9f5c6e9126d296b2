#!/bin/bash
# Compile the reference CaPS hot path (libcaps + cquadpack + libhodlr/Eigen,
# all vendored sources of the reference tree) into oracle/_ref/libcaps_ref.so.
#
# This is test infrastructure only: the shared object is a checker for the
# CPU oracle and a CPU baseline; nothing in the product links it.
#
# The reference's own build (CMake) needs MPI, BLAS and LAPACK, none of which
# this image provides.  The hot path with the reference's default determinant
# algorithm (HODLR, Eigen-based) needs none of them: only the dense-LAPACK
# determinant routines in matrix.c refer to dgetrf_/dpotrf_/dgeqrf_/dgemm_/
# ddot_.  Those symbols stay undefined in the .so (no stand-in is written);
# the library is loaded with lazy binding, so any call into the LAPACK route
# of the reference would fail loudly, while caps_logdetD with DETALG_HODLR
# works.  psd.c (needs dstemr_) is not part of the hot path and is left out.
set -euo pipefail
REF=${CAPS_REFERENCE:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
OUT=$HERE/_ref
OBJ=$OUT/obj
if [ ! -d "$REF/src" ]; then
    echo "reference tree not found at $REF; oracle/_ref not built" >&2
    exit 0
fi
mkdir -p "$OBJ"
CFLAGS="-O3 -std=c99 -fPIC -w -I$REF/src/include -I$REF/src/cquadpack/include -I$REF/src/libhodlr/include -DLIBCAPS_VERSION=\"0.5\""
CXXFLAGS="-O3 -std=c++11 -fPIC -w -DUSE_DOUBLE -I$REF/src/libhodlr/include"
CSRC="bessel.c cache.c integration.c libcaps.c logfac.c matrix.c misc.c plm.c utils.c material.c"
QSRC="dqage.c dqagi.c dqags.c dqext.c dqk15.c dqk15i.c dqk21.c dqk31.c dqk41.c dqk51.c dqk61.c dqsort.c"
HSRC="hodlr.cpp HODLR_Matrix.cpp HODLR_Node.cpp HODLR_Tree.cpp HODLR_Tree_NonSPD.cpp HODLR_Tree_SPD.cpp KDTree.cpp"
pids=()
for f in $CSRC; do gcc $CFLAGS -c "$REF/src/$f" -o "$OBJ/${f%.c}.o" & pids+=($!); done
for f in $QSRC; do gcc $CFLAGS -c "$REF/src/cquadpack/src/$f" -o "$OBJ/q_${f%.c}.o" & pids+=($!); done
for f in $HSRC; do g++ $CXXFLAGS -c "$REF/src/libhodlr/src/$f" -o "$OBJ/h_${f%.cpp}.o" & pids+=($!); done
for p in "${pids[@]}"; do wait "$p"; done
g++ -shared -o "$OUT/libcaps_ref.so" "$OBJ"/*.o -lm
echo "built $OUT/libcaps_ref.so"

# Drop-in check: the reference's block CLI (src/caps_logdetD.c, with the
# material/utility sources it needs for its own option handling) compiled
# against the reference's headers and linked, unchanged, against
# libcaps_amd.so instead of the reference library.
AMD=$HERE/../caps_amd/libcaps_amd.so
if [ -f "$AMD" ]; then
    gcc -O2 -std=c99 -D_GNU_SOURCE -w -I"$REF/src/include" \
        "$REF/src/caps_logdetD.c" "$REF/src/material.c" "$REF/src/utils.c" "$REF/src/misc.c" \
        -L"$(dirname "$AMD")" -lcaps_amd -Wl,-rpath,'$ORIGIN/../../caps_amd' -lm -o "$OUT/caps_logdetD_amd"
    echo "built $OUT/caps_logdetD_amd (reference CLI linked against libcaps_amd.so)"
fi
