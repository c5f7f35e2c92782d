#!/bin/bash
# The CPU test suite (pytest -m "not gpu") with the library's host code and
# the oracle built under AddressSanitizer and UndefinedBehaviorSanitizer.
# CPU only: no test in that selection launches a kernel.
#   bash tests/run_sanitized.sh [pytest args]
set -euo pipefail
ROOT=$(cd "$(dirname "$0")/.." && pwd)
make -s -C "$ROOT/compute-vp9_amd" -j8 sanitize
make -s -C "$ROOT/oracle" sanitize
RT="$(/opt/rocm/llvm/bin/clang -print-resource-dir)/lib/linux/libclang_rt.asan-x86_64.so"
export CVP9_HIP_LIB="$ROOT/compute-vp9_amd/_build_san/libcvp9_hip.so"
export CVP9_ORACLE_LIB="$ROOT/oracle/_build_san/libcvp9_oracle.so"
# leaks: the Python interpreter's own allocations at exit are not ours
export ASAN_OPTIONS=detect_leaks=0:abort_on_error=1:strict_string_checks=1:detect_stack_use_after_return=1
export UBSAN_OPTIONS=halt_on_error=1:print_stacktrace=1
LD_PRELOAD="$RT" python3 -m pytest "$ROOT/tests" -q -m "not gpu" -p no:cacheprovider "$@"
