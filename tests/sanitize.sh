#!/bin/bash
# Runs the CPU test suite against instrumented builds of the oracle and of the
# host side of libzipgpu (test infrastructure; builds go to $OUT, default
# /tmp/zp_san):
#   1. ASan + UBSan: both libraries built with clang, the clang ASan runtime
#      preloaded into Python; every CPU test.
#   2. TSan: the oracle (its threaded loops and the threaded CPU baseline),
#      the tests that drive it.
# Device code is built without sanitizers (-Xarch_host applies them to the
# host pass only).
set -euo pipefail
ROOT=$(cd "$(dirname "$0")/.." && pwd)
OUT=${OUT:-/tmp/zp_san}
CLANG=/opt/rocm/llvm/bin/clang++
RT=$(ls -d /opt/rocm/llvm/lib/clang/*/lib/linux | head -1)
mkdir -p "$OUT/asan" "$OUT/tsan"
rm -f "$OUT"/asan/report.* "$OUT"/tsan/report.*
SAN="-fsanitize=address,undefined -fno-sanitize-recover=undefined -fno-omit-frame-pointer -g -O1"

echo "[san] oracle (ASan+UBSan)"
$CLANG $SAN -std=c++17 -fPIC -shared -o "$OUT/asan/liboracle.so" \
  "$ROOT/oracle/zip_oracle.cpp" "$ROOT/oracle/baseline.cpp" -lpthread

echo "[san] libzipgpu host code (ASan+UBSan)"
HFL="--offload-arch=gfx950 -O1 -g -std=c++17 -fPIC -Xclang -target-feature -Xclang -packed-fp32-ops"
HSAN="-Xarch_host -fsanitize=address -Xarch_host -fsanitize=undefined -Xarch_host -fno-sanitize-recover=undefined -Xarch_host -fno-omit-frame-pointer"
objs=()
for f in "$ROOT"/zinc-plus_amd/csrc/*.hip; do
  o="$OUT/asan/$(basename "${f%.hip}").o"
  /opt/rocm/bin/hipcc $HFL $HSAN -c "$f" -o "$o" &
  objs+=("$o")
done
$CLANG $SAN -std=c++17 -fPIC -D__HIP_PLATFORM_AMD__ -I/opt/rocm/include \
  -c "$ROOT/zinc-plus_amd/csrc/params.cpp" -o "$OUT/asan/params.o" &
wait
/opt/rocm/bin/hipcc --offload-arch=gfx950 -shared -fPIC -o "$OUT/asan/libzipgpu.so" "${objs[@]}" "$OUT/asan/params.o"

echo "[san] CPU tests under ASan+UBSan"
(cd "$ROOT" && ZP_ORACLE_LIB="$OUT/asan/liboracle.so" ZIPGPU_LIB="$OUT/asan/libzipgpu.so" \
  LD_PRELOAD="$RT/libclang_rt.asan-x86_64.so" ASAN_OPTIONS=detect_leaks=0:halt_on_error=1:log_path=$OUT/asan/report \
  UBSAN_OPTIONS=halt_on_error=1:print_stacktrace=1:log_path=$OUT/asan/report \
  python -m pytest tests -q -m "not gpu" -p no:cacheprovider)

echo "[san] oracle (TSan)"
$CLANG -fsanitize=thread -g -O1 -std=c++17 -fPIC -shared -o "$OUT/tsan/liboracle.so" \
  "$ROOT/oracle/zip_oracle.cpp" "$ROOT/oracle/baseline.cpp" -lpthread
(cd "$ROOT" && ZP_ORACLE_LIB="$OUT/tsan/liboracle.so" \
  LD_PRELOAD="$RT/libclang_rt.tsan-x86_64.so" TSAN_OPTIONS=halt_on_error=1:report_signal_unsafe=0:log_path=$OUT/tsan/report \
  python -m pytest tests/test_oracle.py tests/test_baseline.py tests/test_golden.py -q -m "not gpu" \
  -p no:cacheprovider)
echo "[san] clean"
