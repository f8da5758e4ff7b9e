#!/bin/bash
# Regenerates tests/golden/fixture.fun and fixture.g3d from fun_fixture.F90.
set -e
HERE=$(cd "$(dirname "$0")" && pwd)
TMP=$(mktemp -d)
/opt/rocm/bin/amdflang -O0 "$HERE/fun_fixture.F90" -o "$TMP/fun_fixture"
(cd "$HERE" && "$TMP/fun_fixture")
rm -rf "$TMP"
