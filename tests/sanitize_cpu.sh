#!/bin/bash
# Runs the CPU test suite against AddressSanitizer + UndefinedBehaviorSanitizer
# builds of the oracle (oracle/_san) and of the C++ host driver
# (b2r_amd/host/_san: the reference-named shims, the argument checks of
# b2r_main and the on-disk writers).  Host code only: the HIP library is
# loaded as built.  Python itself is not instrumented, so the sanitizer
# runtimes are preloaded; leak checking is off (the interpreter keeps its
# allocations to the end).  Any report ends the run with a non-zero status.
#   tests/sanitize_cpu.sh [extra pytest arguments]
set -euo pipefail
ROOT="$(cd "$(dirname "${BASH_SOURCE[0]}")/.." && pwd)"
make -C "$ROOT/oracle" sanitize
make -C "$ROOT/b2r_amd/host" sanitize
export B2R_ORACLE_LIB="$ROOT/oracle/_san/liboracle.so"
export B2R_HOST_LIB="$ROOT/b2r_amd/host/_san/libb2r_host.so"
export B2R_MAIN="$ROOT/b2r_amd/host/_san/b2r_main"
# alloc_dealloc_mismatch=0: the reference's own heat_3D_cpu_main, compiled into
# oracle/_ref for the comparison tests, allocates with new[] and releases with
# delete (heat-3D.cpp:144, 176-178); this build's code has no such pair.
# Reports go to files (log_path) because pytest captures the test's stderr.
SANLOG="$(mktemp -d)"
export ASAN_OPTIONS="detect_leaks=0:halt_on_error=1:alloc_dealloc_mismatch=0:log_path=$SANLOG/asan"
export UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1:log_path=$SANLOG/ubsan"
export UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1"
PRELOAD="$(gcc -print-file-name=libasan.so) $(gcc -print-file-name=libubsan.so)"
cd "$ROOT"
rc=0
LD_PRELOAD="$PRELOAD" python -m pytest tests -m "not gpu" -x -q -p no:cacheprovider "$@" || rc=$?
for f in "$SANLOG"/*; do
    [ -e "$f" ] || continue
    echo "== sanitizer report $f"
    cat "$f"
    rc=1
done
[ $rc -eq 0 ] && echo "sanitizer run: no reports"
exit $rc
