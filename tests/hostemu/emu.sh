#!/bin/bash
# Run a Python program against the host emulation build under
# AddressSanitizer: tests/hostemu/emu.sh tests/hostemu/run_case.py gubser 0.2 2
set -e
HERE="$(cd "$(dirname "$0")" && pwd)"
export LD_PRELOAD="$(g++ -print-file-name=libasan.so)"
export ASAN_OPTIONS="detect_leaks=0:abort_on_error=1:halt_on_error=1"
export CCAKE_AMD_LIBRARY="$HERE/_build/libccake_emu.so"
exec python3 "$@"
