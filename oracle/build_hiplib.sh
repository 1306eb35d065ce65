#!/bin/bash
# Builds SPILADY's own program (its main loop, initialize, core, check, scale
# and write) from its sources where they lie, with
# spilady_amd/csrc/spilady_hip_bridge.cpp in place of its CPU/CUDA backends,
# linked against spilady_amd/libspilady_hip.so: the drop-in integration a
# SPILADY maintainer would make, built exactly as the bridge's header
# comment describes. Output: oracle/_ref/hiplib/spilady (git-ignored; it is
# run as a check, never shipped as the product).
#
#   oracle/build_hiplib.sh [REF] [extra -D switches...]
#
# The five "#ifdef GPU" lines that call the GPU lifecycle functions are
# widened to "#if defined GPU || defined HIPLIB" as each file is compiled
# (sed into the compiler's stdin: no copy of a source is written anywhere);
# each target line is checked to hold exactly "#ifdef GPU" first.
set -euo pipefail
HERE="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
REPO="$(dirname "$HERE")"
REF="${1:-/root/reference}"
shift || true
EXTRA=("$@")
OUT="${SLD_HIPLIB_OUT:-$HERE/_ref/hiplib}"
CXX="${CXX:-g++}"
LIB="$REPO/spilady_amd/libspilady_hip.so"
[ -f "$REF/spilady.cpp" ] || { echo "no SPILADY sources at $REF" >&2; exit 2; }
[ -f "$LIB" ] || { echo "$LIB is not built (make -C spilady_amd/csrc)" >&2; exit 2; }
FLAGS=(-O2 -fopenmp -w -DHIPLIB -DSLDHL "${EXTRA[@]}" -include "$REPO/spilady_amd/csrc/spilady_hip_bridge.h"
       -I"$REPO/include" -I"$REPO/spilady_amd/csrc")
declare -A WIDEN=([initial.cpp]="67" [write_config.cpp]="35 137" [write_vsim.cpp]="37" [free_memory.cpp]="41")
mkdir -p "$OUT/obj"

compile() {
    local src="$1" name
    name="$(basename "$src")"
    local obj="$OUT/obj/${name%.cpp}.o"
    if [ -n "${WIDEN[$name]:-}" ]; then
        local sedargs=()
        for l in ${WIDEN[$name]}; do
            if ! sed -n "${l}p" "$src" | grep -qE '^[[:space:]]*#ifdef GPU[[:space:]]*$'; then
                echo "$name:$l is not an '#ifdef GPU' line" >&2
                return 1
            fi
            sedargs+=(-e "${l}s/#ifdef GPU/#if defined GPU || defined HIPLIB/")
        done
        sed "${sedargs[@]}" "$src" | "$CXX" "${FLAGS[@]}" -I"$REF" -x c++ -c - -o "$obj"
    else
        "$CXX" "${FLAGS[@]}" -c "$src" -o "$obj"
    fi
}
pids=()
for src in "$REF"/*.cpp; do
    compile "$src" &
    pids+=($!)
    if [ "${#pids[@]}" -ge 8 ]; then wait "${pids[0]}"; pids=("${pids[@]:1}"); fi
done
for p in "${pids[@]}"; do wait "$p"; done
# the bridge's own definitions must each match a prototype SPILADY declares
"$CXX" "${FLAGS[@]}" -I"$REF" -Werror=missing-declarations -c "$REPO/spilady_amd/csrc/spilady_hip_bridge.cpp" \
    -o "$OUT/obj/zz_spilady_hip_bridge.o"
"$CXX" -fopenmp -o "$OUT/spilady" "$OUT"/obj/*.o -L"$REPO/spilady_amd" -lspilady_hip \
    -Wl,-rpath,'$ORIGIN/../../../spilady_amd' -Wl,-rpath,"$REPO/spilady_amd"
echo "built $OUT/spilady"
