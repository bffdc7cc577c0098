#!/bin/sh
# Regenerates tests/golden/philox_kat.json from PyTorch's Philox header (CPU only).
set -e
HERE=$(cd "$(dirname "$0")" && pwd)
TORCH_INC=$(python3 -c 'import os, torch; print(os.path.join(os.path.dirname(torch.__file__), "include"))')
OUT=${TMPDIR:-/tmp}/gen_philox_kat
g++ -std=c++17 -O1 -I"$TORCH_INC" "$HERE/gen_philox_kat.cpp" -o "$OUT"
"$OUT" > "$HERE/philox_kat.json"
