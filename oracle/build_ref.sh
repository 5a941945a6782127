#!/usr/bin/env bash
# Build oracle/_ref/libga_ref.so (the reference's own pairwise kernels and API)
# and oracle/_ref/gene_qc (its gene QC tool on that API), compiled for the CPU
# with the reference's own HIP CPU shim, from the sources
# where they lie under the reference checkout.  TEST INFRASTRUCTURE ONLY.
#
#   oracle/build_ref.sh [REFERENCE_ROOT]     (default /root/reference)
#
# Exits 0 without building when the reference checkout is absent (the GPU box
# has no reference tree; it uses the prebuilt library if one travelled).
set -euo pipefail
REF="${1:-/root/reference}"
HERE="$(cd "$(dirname "$0")" && pwd)"
OUT="$HERE/_ref"
if [ ! -f "$REF/src/api/api.cpp" ] || [ ! -f "$REF/tests/parity/hip_cpu_shim/hip/hip_runtime.h" ]; then
    echo "build_ref: reference tree not found at $REF; skipping" >&2
    exit 0
fi
mkdir -p "$OUT"
g++ -O2 -std=c++17 -pthread -fPIC -shared -ffp-contract=off \
    -DGENOALIGNER_HIP_SHIM \
    -I "$REF/tests/parity/hip_cpu_shim" -I "$REF" -I "$REF/include" \
    "$HERE/ref/ref_driver.cpp" \
    -o "$OUT/libga_ref.so.tmp"
mv "$OUT/libga_ref.so.tmp" "$OUT/libga_ref.so"
echo "build_ref: wrote $OUT/libga_ref.so"
# the reference's gene QC tool (tools/gene_qc.cpp) over the same API build
g++ -O2 -std=c++17 -pthread -ffp-contract=off \
    -DGENOALIGNER_HIP_SHIM \
    -I "$REF/tests/parity/hip_cpu_shim" -I "$REF" -I "$REF/include" \
    "$REF/tools/gene_qc.cpp" "$HERE/ref/ref_driver.cpp" \
    -o "$OUT/gene_qc.tmp"
mv "$OUT/gene_qc.tmp" "$OUT/gene_qc"
echo "build_ref: wrote $OUT/gene_qc"
# the reference's MSA building blocks (src/msa/msa_ref.cpp) and its
# profile-profile kernel body over the same shim
g++ -O2 -std=c++17 -pthread -fPIC -shared -ffp-contract=off \
    -DGENOALIGNER_HIP_SHIM \
    -I "$REF/tests/parity/hip_cpu_shim" -I "$REF" -I "$REF/include" \
    "$HERE/ref/ref_msa_driver.cpp" "$REF/src/msa/msa_ref.cpp" \
    -o "$OUT/libga_ref_msa.so.tmp"
mv "$OUT/libga_ref_msa.so.tmp" "$OUT/libga_ref_msa.so"
echo "build_ref: wrote $OUT/libga_ref_msa.so"
