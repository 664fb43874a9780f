#!/usr/bin/env bash
# Compiles the reference GraphWalker apps from their own sources (where they
# lie under /root/reference) into oracle/_ref/, with the flags the reference
# Makefile lists as its optimised build (Makefile:5: -O3 -fopenmp).  Only the
# apps that compile against the current reference headers are built.
# Outputs go to oracle/_ref/ only; nothing from the reference is copied.
set -euo pipefail
REF=${GW_REFERENCE:-/root/reference}
OUT="$(cd "$(dirname "$0")" && pwd)/_ref"
if [ ! -d "$REF/src" ]; then
    echo "build_ref: $REF/src not present; skipping" >&2
    exit 0
fi
mkdir -p "$OUT"
for app in rawrandomwalks msppr simrank graphlet rwdomination kk_ppr; do
    src="$REF/apps/$app.cpp"
    [ -f "$src" ] || continue
    if [ ! -x "$OUT/$app" ] || [ "$src" -nt "$OUT/$app" ]; then
        g++ -O3 -fopenmp -w -I "$REF/src" "$src" -o "$OUT/$app" -lz
    fi
done
# A test-infrastructure app of this repository on the reference API: the
# reference engine's single-source PPR visit vector (oracle/ref_visit_counts.cpp).
HERE="$(cd "$(dirname "$0")" && pwd)"
if [ ! -x "$OUT/visit_counts" ] || [ "$HERE/ref_visit_counts.cpp" -nt "$OUT/visit_counts" ]; then
    g++ -O3 -fopenmp -w -I "$REF/src" "$HERE/ref_visit_counts.cpp" -o "$OUT/visit_counts" -lz
fi
# ... and its batched multi-source PPR on a list of sources (oracle/ref_msppr_sources.cpp).
if [ ! -x "$OUT/msppr_sources" ] || [ "$HERE/ref_msppr_sources.cpp" -nt "$OUT/msppr_sources" ]; then
    g++ -O3 -fopenmp -w -I "$REF/src" "$HERE/ref_msppr_sources.cpp" -o "$OUT/msppr_sources" -lz
fi
echo "build_ref: built $(ls "$OUT" | tr '\n' ' ')"
