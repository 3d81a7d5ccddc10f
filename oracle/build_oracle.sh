#!/bin/bash
# build_oracle.sh -- compile the C restatement (oracle/qs3_oracle.c) into
# oracle/liboracle_qs3.so.  Test infrastructure only.
set -euo pipefail
HERE="$(cd "$(dirname "$0")" && pwd)"
gcc -O2 -fopenmp -fPIC -shared -std=c11 -Wall -Wno-unused-variable \
  -o "$HERE/liboracle_qs3.so" "$HERE/qs3_oracle.c" -lm
echo "built $HERE/liboracle_qs3.so"
