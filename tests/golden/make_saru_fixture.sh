#!/usr/bin/env bash
# Regenerates tests/golden/saru_ref_vectors.txt from the reference's own
# Saru header (include/saruprng.hpp), built by oracle/build_ref.sh.  Run
# where the reference tree is present; the GPU tests only read the committed
# fixture.
set -euo pipefail
HERE=$(cd "$(dirname "$0")" && pwd)
ROOT=$(cd "$HERE/../.." && pwd)
"$ROOT/oracle/build_ref.sh"
"$ROOT/oracle/_ref/saru_vectors_fastmath" > "$HERE/saru_ref_vectors.txt"
# the IEEE-strict build must print the same vectors
"$ROOT/oracle/_ref/saru_vectors_strict" | cmp - "$HERE/saru_ref_vectors.txt"
echo "wrote $HERE/saru_ref_vectors.txt"
