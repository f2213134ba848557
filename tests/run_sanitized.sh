#!/bin/bash
# Runs the CPU test suite on AddressSanitizer + UndefinedBehaviorSanitizer
# builds of the engine's host code and of the oracle (no GPU needed):
#
#   python -m xmcml_amd.build --sanitize && bash tests/run_sanitized.sh [pytest args]
#
# Both are built by clang (hipcc for the library, host side only) and share
# clang's sanitizer runtime, which is preloaded because the interpreter is not
# instrumented; leak checking is off (the HIP runtime and Python keep memory
# until exit).
set -euo pipefail
ROOT="$(cd "$(dirname "$0")/.." && pwd)"
cd "$ROOT"
export LD_PRELOAD="$(ls /opt/rocm/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so | head -1)"
export ASAN_OPTIONS="detect_leaks=0:halt_on_error=1:abort_on_error=1:detect_odr_violation=0"
export UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1"
export XMC_LIBRARY="$ROOT/xmcml_amd/libxmcml_amd_san.so"
export XMC_ORACLE_LIBRARY="$ROOT/oracle/liboracle_san.so"
exec python -m pytest tests -m "not gpu" -q -p no:cacheprovider "$@"
