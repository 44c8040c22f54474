#!/bin/bash
# Runs the CPU test suite against the host code built with AddressSanitizer and UndefinedBehaviorSanitizer.
#
# In a scratch copy of the tree (the product's own build is not touched), the g++-compiled host part of the library
# (the file formats, the tag-similarity graph, the generators and recounts: cuda_graph_clustering_amd/csrc/*.cpp),
# the cc_gpu command line and the C oracle (oracle/cc_oracle.c) are rebuilt with -fsanitize=address,undefined and
# -fno-sanitize-recover=all; the HIP translation units keep the objects already built in the tree (their device code
# cannot be sanitized, and mixing clang's and gcc's sanitizer runtimes in one process is not possible). Then
# `pytest -m "not gpu"` runs there with gcc's ASan runtime preloaded into Python. Every loader test feeds malformed
# and random files, so the parsers run under the sanitizers on untrusted input.
#
# Usage: tests/sanitize.sh [extra pytest arguments]      (needs the tree built first: python -c "import
# __graft_entry__ as g; g.build()")
set -euo pipefail
SRC=$(cd "$(dirname "$0")/.." && pwd)
WORK=$(mktemp -d /tmp/cc_sanitize.XXXXXX)
trap 'rm -rf "$WORK"' EXIT
tar -C "$SRC" --exclude=./.git --exclude=./_work --exclude=./gpurun_out -cf - . | tar -C "$WORK" -xpf -
SAN="-fsanitize=address,undefined -fno-sanitize-recover=all -fno-omit-frame-pointer -g"
UBSAN_DIR=$(dirname "$(g++ -print-file-name=libubsan.so)")
PKG="$WORK/cuda_graph_clustering_amd"
rm -f "$PKG/build/cc_host.o" "$PKG/build/cc_graph_io.o" "$PKG/build/cc_cli.o" "$PKG/lib/libcc_hip.so" "$PKG/bin/cc_gpu"
make -C "$PKG" -j8 CXXFLAGS="-O1 -std=c++17 -fPIC -Wall -pthread $SAN" \
  LIB_EXTRA="-L$UBSAN_DIR -lubsan" CLI_EXTRA="$SAN"
make -C "$WORK/oracle" -B liboracle.so CFLAGS="-O1 -std=c11 -fPIC -pthread -ffp-contract=off -Wall -Wextra $SAN"
cd "$WORK"
# libstdc++ is preloaded right after the ASan runtime: Python itself does not link it, and without it at the
# runtime's start the runtime cannot find the real __cxa_throw behind its interceptor.
LD_PRELOAD="$(g++ -print-file-name=libasan.so) $(g++ -print-file-name=libstdc++.so)" \
  ASAN_OPTIONS="detect_leaks=0:abort_on_error=1" \
  UBSAN_OPTIONS="print_stacktrace=1" python -m pytest tests -m "not gpu" -x -q -p no:cacheprovider "$@"
