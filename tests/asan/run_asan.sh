#!/bin/bash
# Host AddressSanitizer runs, CPU only (no GPU is touched):
#   1. the oracle built with g++ -fsanitize=address,undefined, under its CPU tests;
#   2. the HIP library's host code built with clang's -fsanitize=address (device
#      code unchanged), under the ABI, argument-validation, PLY and Delaunay tests.
# Usage: tests/asan/run_asan.sh   (from anywhere; builds both into _build/asan/)
set -euo pipefail
ROOT=$(cd "$(dirname "$0")/../.." && pwd)
cd "$ROOT"
make -s -C oracle asan
python3 blade-volume_amd/bv/build.py --asan > /dev/null
CLANG_ASAN=$(ls /opt/rocm/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so | head -1)
GCC_ASAN=$(gcc -print-file-name=libasan.so)
# leaks: CPython and the HIP runtime keep allocations to exit by design
export ASAN_OPTIONS=detect_leaks=0:protect_shadow_gap=0:detect_odr_violation=0:abort_on_error=0
export UBSAN_OPTIONS=print_stacktrace=1:halt_on_error=1
echo "== canary: the sanitizer build of the oracle reports an out-of-bounds write"
if LD_PRELOAD="$GCC_ASAN" BVO_LIBRARY="$ROOT/_build/asan/liboracle_asan.so" python3 tests/asan/canary.py \
    2> "$ROOT/_build/asan/canary.log"; then
  echo "canary: no sanitizer report"; exit 1
fi
grep -q "heap-buffer-overflow" "$ROOT/_build/asan/canary.log" && echo "canary: heap-buffer-overflow reported"
echo "== oracle (g++ ASan + UBSan)"
LD_PRELOAD="$GCC_ASAN" BVO_LIBRARY="$ROOT/_build/asan/liboracle_asan.so" python3 -m pytest -q -m "not gpu" \
  -p no:cacheprovider tests/test_oracle.py tests/test_oracle_adjoint.py tests/test_splat_oracle.py \
  tests/test_gaussian_oracle.py tests/test_relight.py tests/test_objects_oracle.py tests/test_graph_oracle.py \
  tests/test_splat_jacobian_oracle.py tests/test_recorder_rays_oracle.py
echo "== HIP library host code (clang ASan)"
LD_PRELOAD="$CLANG_ASAN" BV_LIBRARY="$ROOT/_build/asan/libbvhip_asan.so" python3 -m pytest -q -m "not gpu" \
  -p no:cacheprovider tests/test_abi.py tests/test_delaunay.py tests/test_ply.py
