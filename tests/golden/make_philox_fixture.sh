#!/bin/bash
# Regenerates philox_rocrand.json from rocRAND's host engine (ROCm headers).
set -euo pipefail
cd "$(dirname "$0")"
tmp=$(mktemp -d)
hipcc -O1 -x hip --offload-host-only -o "$tmp/gen" philox_rocrand_vectors.cpp
"$tmp/gen" > philox_rocrand.json
rm -rf "$tmp"
