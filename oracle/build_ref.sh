#!/usr/bin/env bash
# Builds test binaries in oracle/_ref/ from the reference's own sources
# where they lie (only self-contained files: include/saruprng.hpp needs
# nothing but <cmath>).  The rest of the reference path needs the FDPS
# framework, which is an empty submodule in the reference tree, so it is not
# built here.  Outputs go only to oracle/_ref/ (git-ignored).
#
# Flags follow the reference's Makefile (g++ -O3 -std=c++11 -ffast-math
# -funroll-loops, Makefile:10-23) for one build, and IEEE-strict -O2 for the
# other, so that the vectors show whether -ffast-math moves nrml().
set -euo pipefail
REF=${REF_ROOT:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
OUT="$HERE/_ref"
if [ ! -f "$REF/include/saruprng.hpp" ]; then
  echo "build_ref.sh: $REF/include/saruprng.hpp not found; skipping" >&2
  exit 0
fi
mkdir -p "$OUT"
g++ -O3 -std=c++11 -ffast-math -funroll-loops -I"$REF/include" \
    "$HERE/ref_saru_vectors.cpp" -o "$OUT/saru_vectors_fastmath"
g++ -O2 -std=c++11 -I"$REF/include" \
    "$HERE/ref_saru_vectors.cpp" -o "$OUT/saru_vectors_strict"
echo "built $OUT/saru_vectors_fastmath $OUT/saru_vectors_strict"
