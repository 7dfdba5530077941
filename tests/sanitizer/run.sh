#!/bin/bash
# AddressSanitizer + UndefinedBehaviorSanitizer build of the host code and the CPU tests that drive it.
#
# Instrumented (clang, one sanitizer runtime for the whole process):
#   oracle/fp_oracle.cpp                  the CPU oracle
#   firepony_amd/csrc/synth.cpp           the workload generator
#   firepony_amd/csrc/loaders.cpp         FASTA / VCF / SAM / BAM parsers (untrusted input bytes)
#   the engine library, host side only    engine.cpp, report.cpp (device code is built as usual)
# Then the loader, oracle, golden, window-model and ABI tests run against those libraries in a copy
# of the tree under tests/sanitizer/_build/tree; any sanitizer report fails the run.
#
#   bash tests/sanitizer/run.sh
set -euo pipefail
root="$(cd "$(dirname "$0")/../.." && pwd)"
out="$root/tests/sanitizer/_build"
tree="$out/tree"
clang="/opt/rocm/llvm/bin/clang++"
hipcc="/opt/rocm/bin/hipcc"
rt="$(ls /opt/rocm/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so | head -1)"
san="-fsanitize=address,undefined -fno-sanitize-recover=undefined -fno-omit-frame-pointer -shared-libasan -g -O1"

rm -rf "$tree"
mkdir -p "$tree"
(cd "$root" && tar --exclude='_build' --exclude='_lib' --exclude='__pycache__' --exclude='*.so' \
    -cf - firepony_amd oracle tests include) | (cd "$tree" && tar -xf -)
mkdir -p "$tree/firepony_amd/_lib" "$tree/oracle/_build"

echo "+ oracle, generator, loaders (clang, ASan+UBSan)"
$clang $san -std=c++17 -fPIC -shared -pthread -ffp-contract=off -fno-fast-math \
    "$root/oracle/fp_oracle.cpp" -o "$tree/oracle/_build/libfp_oracle.so"
$clang $san -std=c++17 -fPIC -shared -pthread "$root/firepony_amd/csrc/synth.cpp" \
    -o "$tree/firepony_amd/_lib/libfpsynth.so"
$clang $san -std=c++17 -fPIC -shared -pthread -I "$root/include" "$root/firepony_amd/csrc/loaders.cpp" -lz \
    -o "$tree/firepony_amd/_lib/libfirepony_amd_io.so"

echo "+ engine library, host code instrumented"
cd "$root/firepony_amd/csrc"
$hipcc -x hip -O1 -g -std=c++17 -fPIC -shared --offload-arch=gfx950 -ffp-contract=off \
    -Xclang -target-feature -Xclang -packed-fp32-ops \
    -Xarch_host -fsanitize=address -Xarch_host -fsanitize=undefined -Xarch_host -fno-sanitize-recover=undefined \
    -Xarch_host -shared-libasan -Xarch_host -fno-omit-frame-pointer \
    -I "$root/include" -I "$root/firepony_amd/csrc" \
    engine.cpp report.cpp k_prep.hip k_baq.hip k_cov.hip k_sort.hip k_table.hip k_post.hip \
    -o "$tree/firepony_amd/_lib/libfirepony_amd.so" 2>&1 | grep -v "packed-fp32-ops\|warning generated" || true
test -s "$tree/firepony_amd/_lib/libfirepony_amd.so"

echo "+ CPU tests under the sanitizers"
cd "$tree"
export ASAN_OPTIONS="detect_leaks=0:abort_on_error=1:halt_on_error=1:strict_string_checks=1"
export UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1"
LD_PRELOAD="$rt" python3 -m pytest -x -q -p no:cacheprovider -m "not gpu" \
    tests/test_loaders.py tests/test_oracle_pins.py tests/test_golden.py tests/test_oracle_windows.py \
    tests/test_abi.py tests/test_report_compat.py 2>&1 | tee "$out/run.log" | tail -5
