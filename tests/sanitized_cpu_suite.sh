#!/bin/bash
# The CPU test suite with the host C++ of the library (input, set-up, writers,
# C-ABI) and the oracle built under AddressSanitizer + UndefinedBehaviorSanitizer.
# No GPU is needed: device calls fail cleanly with SPH_ERR_DEVICE there.
set -e
ROOT="$(cd "$(dirname "$0")/.." && pwd)"
PKG="$ROOT/sph-magnetic-particles-oldroyd-b-in-cuda_amd"
python3 "$PKG/build.py" --sanitize
make -s -C "$ROOT/oracle" asan
export SPH_LIB="$PKG/libsph_amd_asan.so"
export SPH_ORACLE_LIB="$ROOT/oracle/liboracle_asan.so"
export LD_PRELOAD="$(g++ -print-file-name=libasan.so):$(g++ -print-file-name=libubsan.so)"
export ASAN_OPTIONS=detect_leaks=0:abort_on_error=1
export UBSAN_OPTIONS=print_stacktrace=1:halt_on_error=1
cd "$ROOT"
python3 -m pytest tests -q -m "not gpu" -p no:cacheprovider "$@"
