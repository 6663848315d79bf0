#!/usr/bin/env bash
# Compile the reference's own CPU path (source/main.cu, modes 4/5) into
# oracle/_ref/ so that tests can pin the C oracle against it bit for bit.
#
# The reference's main.cu also holds CUDA host code (Thrust, kernel
# launches), so it is not compiled whole: g++ reads only the line ranges that
# hold the standard-library includes and the pure C++ CPU functions, straight
# from the reference tree, followed by oracle/ref_driver.cpp.  The -D values
# are the reference's Python-library flags (source/Makefile:47-91); the
# compiler flags are its host flags (source/Makefile:13: -std=c++14 -O3
# -fopenmp).  Nothing is written outside oracle/_ref/, and no reference
# source is copied into this repository.
#
# Usage: oracle/build_ref.sh [REFERENCE_ROOT]   (default /root/reference)
set -euo pipefail

REF="${1:-/root/reference}"
SRC="$REF/source/main.cu"
HERE="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
OUT="$HERE/_ref"

if [ ! -f "$SRC" ]; then
    echo "build_ref: $SRC not found; reference CPU build skipped" >&2
    exit 0
fi
mkdir -p "$OUT"

# Line ranges of source/main.cu (checked by the sentinel greps below):
#   1-15     standard-library and OpenMP includes
#   50-90    function prototypes and `using namespace std;`
#   468-510  computeObjectRanges
#   1295-1777 the CPU path (normalization, modes 4/5 drivers, computePeriod,
#            the standard and generalized inner loops, updateYerrorfactor)
RANGES='1,15p;50,90p;468,510p;1295,1777p'
sed -n '468p' "$SRC" | grep -q '^void computeObjectRanges' || { echo "build_ref: unexpected main.cu layout (468)" >&2; exit 1; }
sed -n '1295p' "$SRC" | grep -q '^double computeStandardDevSquared' || { echo "build_ref: unexpected main.cu layout (1295)" >&2; exit 1; }
sed -n '1701p' "$SRC" | grep -q '^void lombscarglecpuinnerloopAstroPy' || { echo "build_ref: unexpected main.cu layout (1701)" >&2; exit 1; }
sed -n '90p' "$SRC" | grep -q '^using namespace std;' || { echo "build_ref: unexpected main.cu layout (90)" >&2; exit 1; }

build() {  # name error dtype normalize threads
    local name="$1" err="$2" dt="$3" norm="$4" thr="$5"
    { sed -n "$RANGES" "$SRC"; cat "$HERE/ref_driver.cpp"; } |
        g++ -x c++ -std=c++14 -W -O3 -fopenmp -fPIC -shared -w \
            -I "$REF/source" -include "$REF/source/structs.h" \
            -DPYTHON -DERROR="$err" -DDTYPE="$dt" -DNORMALIZEPGRAM="$norm" \
            -DNTHREADSCPU="$thr" -DNUMGPU=1 -DPRINTPERIODS=0 -DPRINTPGRAM=0 \
            -DBLOCKSIZE=512 -DPINNED=1 -DRETURNPGRAM=1 -DSIZEPINNEDBUFFERMIB=8 \
            -DSHMEM=0 -DNSTREAMS=3 -DALPHA=0.35 \
            - -o "$OUT/$name.so"
}

for dt in float double; do
    for thr in 1 16; do
        build "libref_noerror${dt}_t${thr}" 0 "$dt" 1 "$thr" &
        build "libref_noerror${dt}_raw_t${thr}" 0 "$dt" 0 "$thr" &
        build "libref_error${dt}_t${thr}" 1 "$dt" 1 "$thr" &
    done
    wait
done
echo "build_ref: built $(ls "$OUT"/*.so | wc -l) libraries in $OUT"
