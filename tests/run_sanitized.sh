#!/bin/bash
# The CPU test suite (everything not marked gpu) against host-side
# AddressSanitizer + UndefinedBehaviorSanitizer builds of libzs_hip.so (device code
# is not instrumented) and of the oracle: the wire-format parsers and verifiers on
# valid, tampered, truncated and randomly corrupted proofs, the channel, the C ABI's
# argument checks.  No GPU is used.  Usage: tests/run_sanitized.sh [pytest args]
set -euo pipefail
ROOT=$(cd "$(dirname "$0")/.." && pwd)
python3 "$ROOT/zig-stark_amd/build.py" --sanitize > /dev/null
make -s -C "$ROOT/oracle" san
RT=$(/opt/rocm/llvm/bin/clang -print-file-name=libclang_rt.asan-x86_64.so)
export ZS_HIP_LIBRARY="$ROOT/zig-stark_amd/_build_san/libzs_hip_san.so"
export ZS_ORACLE_LIBRARY="$ROOT/oracle/_build_san/liboracle_san.so"
# the interpreter itself is not instrumented: leak reports would be its own
export ASAN_OPTIONS=detect_leaks=0:abort_on_error=1:strict_string_checks=1:detect_stack_use_after_return=1
export UBSAN_OPTIONS=print_stacktrace=1:halt_on_error=1
# the ASan runtime goes first; anything already preloaded stays (and is recorded, so that
# tests starting their own sanitized programs give them the original list)
export ZS_BASE_PRELOAD="${LD_PRELOAD:-}"
LD_PRELOAD="$RT${LD_PRELOAD:+:$LD_PRELOAD}" python3 -m pytest "$ROOT/tests" -q -m "not gpu" -p no:cacheprovider "$@"
