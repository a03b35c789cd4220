#!/bin/bash
# Builds a program that sigflowc wrote with --emit=hip:
#   gpu-rf-pipeline_amd/sigflow/build_program.sh PROGRAM.cpp OUTPUT
# It links the pipeline runtime (libsigflow_rt.a) and librfhip.so of this package, found through
# a run path relative to OUTPUT, so a tree that moves as a whole keeps working. Build the
# package first: make -C gpu-rf-pipeline_amd. CXX overrides the compiler (default g++).
set -euo pipefail
if [ $# -ne 2 ]; then
    echo "usage: $0 PROGRAM.cpp OUTPUT" >&2
    exit 2
fi
pkg=$(cd "$(dirname "$0")/.." && pwd)
for f in librfhip.so libsigflow_rt.a; do
    if [ ! -f "$pkg/$f" ]; then
        echo "$0: $pkg/$f is not built (make -C $pkg)" >&2
        exit 2
    fi
done
out_dir=$(cd "$(dirname "$2")" && pwd)
rel=$(realpath --relative-to="$out_dir" "$pkg")
"${CXX:-g++}" -O2 -std=c++17 -Wall -I"$pkg/../include" -I"$pkg/host" "$1" -o "$2" \
    "$pkg/libsigflow_rt.a" -L"$pkg" -lrfhip -Wl,-rpath,"\$ORIGIN/$rel"
