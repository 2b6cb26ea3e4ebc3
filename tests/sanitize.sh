#!/bin/bash
# The CPU test suite with the host code of libvdock and the oracle built under AddressSanitizer
# and UndefinedBehaviorSanitizer (clang; device code unchanged).  Python itself is not
# instrumented, so the sanitizer runtime is preloaded; leak checking is off (the interpreter's
# own allocations would swamp it).  Extra arguments go to pytest.
set -e
ROOT="$(cd "$(dirname "$0")/.." && pwd)"
make -s -C "$ROOT/vina-cuda_amd" -j8 san
make -s -C "$ROOT/oracle" san
RT="$(ls /opt/rocm/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so | head -1)"
cd "$ROOT"
LD_PRELOAD="$RT" ASAN_OPTIONS=detect_leaks=0 UBSAN_OPTIONS=print_stacktrace=1 \
  VDOCK_LIBRARY="$ROOT/vina-cuda_amd/_build_san/libvdock.so" ORACLE_LIBRARY="$ROOT/oracle/_build_san/liboracle.so" \
  python -m pytest tests -q -m "not gpu" -p no:cacheprovider "$@"
