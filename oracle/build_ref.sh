#!/bin/bash
# Builds oracle/_ref/libtroute_ref.so from the reference's own FP64 host
# routines, where they lie under the reference tree.  The .cu files cannot be
# compiled whole here (no nvcc, no CUDA headers), but the host routines are
# self-contained C++ using <cmath>/<vector> only.  Their line ranges are cut
# into a temporary translation unit OUTSIDE this repository (nothing of the
# reference source is kept), our C driver is appended, and g++ builds a .so
# into oracle/_ref/ (git-ignored).  Compiled exactly like the oracle:
# -O2 -ffp-contract=off.
#
#   wavefront/wavefront_mc.cu : 20-29 (std includes, skipping cuda_runtime.h at 30)
#                               41-44 (ChParams, float)
#                               200-337 (hgeo_inline_d, mc_solve_d)
#                               495-564 (cpu_wavefront_fp64)
#   the same file, FP32 rule  : 20-29, 41-44, 51-193 (hgeo_inline, mc_solve),
#                               413-491 (cpu_wavefront) -> libtroute_ref_f32.so.
#                               These routines are __host__ __device__, which
#                               g++ does not know: hipcc compiles them for the
#                               host only (--offload-host-only, no device code),
#                               calling the host libm's powf/sqrtf as the
#                               reference's CPU build does.
#   troute_mc_final.cu        : 20-23 (includes, skipping cuda_runtime.h at 24)
#                               26-28 (ChParams, double)
#                               167-252 (cpu_mc), 254-274 (gen)
#   wavefront/linear_mc.cu    : 27-35 (std includes, skipping cuda_runtime.h at 36)
#                               238-289 (cpu_linear_mc_fp64)
#   wavefront/data_parsers/parse_qlat_fast_cpp.cpp: the whole program (no CUDA),
#                               built as oracle/_ref/parse_qlat_fast
#   wavefront/linear_mc_matrix.cu: 26-35 (std includes, skipping cuda_runtime.h
#                               and cusparse.h at 36-37), 55 (ChParams, float),
#                               58-84 (compute_mc_coeffs_cpu), 111-229 (Matrix,
#                               build_propagation_matrices)
#   wavefront/linear_mc_matrix_secant.cu: 23-32 (std includes, skipping
#                               cuda_runtime.h and cusparse.h at 33-34), 52
#                               (ChParams, float), 57-195 (mc_secant_solve_cpu)
#   ../owp_extended_kernels.cu: 13-16 (std includes, skipping cuda_runtime.h at 17)
#                               19-20 (MAX_NODES), 28-32 (DiffReachParams)
#                               35-94 (cpu_diffusive_tridiag), 347-362 (gen_diff_data)
set -euo pipefail
REF=${TROUTE_REF_DIR:-/root/reference/water-prediction/t-route}
HERE=$(cd "$(dirname "$0")" && pwd)
OUT="$HERE/_ref"
if [ ! -f "$REF/wavefront/wavefront_mc.cu" ] || [ ! -f "$REF/troute_mc_final.cu" ] || \
   [ ! -f "$REF/wavefront/linear_mc.cu" ] || [ ! -f "$REF/../owp_extended_kernels.cu" ] || \
   [ ! -f "$REF/wavefront/linear_mc_matrix.cu" ] || [ ! -f "$REF/wavefront/linear_mc_matrix_secant.cu" ]; then
    echo "build_ref: reference sources not present; oracle/_ref not rebuilt"
    exit 0
fi
TMP=$(mktemp -d /tmp/trmc_ref.XXXXXX)
trap 'rm -rf "$TMP"' EXIT
{ sed -n '20,29p;41,44p;200,337p;495,564p' "$REF/wavefront/wavefront_mc.cu"
  cat "$HERE/ref_driver_wavefront.inc"; } > "$TMP/ref_wavefront.cpp"
{ sed -n '20,29p;41,44p;51,193p;413,491p' "$REF/wavefront/wavefront_mc.cu"
  cat "$HERE/ref_driver_wavefront_f32.inc"; } > "$TMP/ref_wavefront_f32.cpp"
{ sed -n '20,23p;26,28p;167,252p;254,274p' "$REF/troute_mc_final.cu"
  cat "$HERE/ref_driver_mcfinal.inc"; } > "$TMP/ref_mcfinal.cpp"
{ sed -n '27,35p;238,289p' "$REF/wavefront/linear_mc.cu"
  cat "$HERE/ref_driver_linear.inc"; } > "$TMP/ref_linear.cpp"
{ sed -n '26,35p;55p;58,84p;111,229p' "$REF/wavefront/linear_mc_matrix.cu"
  cat "$HERE/ref_driver_matrix.inc"; } > "$TMP/ref_matrix.cpp"
{ sed -n '23,32p;52p;57,195p' "$REF/wavefront/linear_mc_matrix_secant.cu"
  cat "$HERE/ref_driver_matrix_secant.inc"; } > "$TMP/ref_matrix_secant.cpp"
{ sed -n '13,16p;19,20p;28,32p;35,94p;347,362p' "$REF/../owp_extended_kernels.cu"
  cat "$HERE/ref_driver_diffusive.inc"; } > "$TMP/ref_diffusive.cpp"
# the cut must land on the routines it names (guards against line drift)
grep -q "static inline double mc_solve_d(" "$TMP/ref_wavefront.cpp"
grep -q "static void cpu_wavefront_fp64(" "$TMP/ref_wavefront.cpp"
grep -q "inline float mc_solve(" "$TMP/ref_wavefront_f32.cpp"
grep -q "static void cpu_wavefront(" "$TMP/ref_wavefront_f32.cpp"
grep -q "^void cpu_mc(" "$TMP/ref_mcfinal.cpp"
grep -q "^void gen(" "$TMP/ref_mcfinal.cpp"
grep -q "static void cpu_linear_mc_fp64(" "$TMP/ref_linear.cpp"
grep -q "^void cpu_diffusive_tridiag(" "$TMP/ref_diffusive.cpp"
grep -q "static void compute_mc_coeffs_cpu(" "$TMP/ref_matrix.cpp"
grep -q "static void build_propagation_matrices(" "$TMP/ref_matrix.cpp"
grep -q "static void mc_secant_solve_cpu(" "$TMP/ref_matrix_secant.cpp"
grep -q "^void gen_diff_data(" "$TMP/ref_diffusive.cpp"
if grep -q "cuda\|cusparse" "$TMP/ref_wavefront.cpp" "$TMP/ref_wavefront_f32.cpp" "$TMP/ref_mcfinal.cpp" "$TMP/ref_linear.cpp" \
        "$TMP/ref_diffusive.cpp" "$TMP/ref_matrix.cpp" "$TMP/ref_matrix_secant.cpp"; then
    echo "build_ref: extracted ranges still reference CUDA" >&2; exit 1
fi
mkdir -p "$OUT"
g++ -O2 -ffp-contract=off -fPIC -shared -std=c++17 -w \
    -o "$OUT/libtroute_ref.so" "$TMP/ref_wavefront.cpp" "$TMP/ref_mcfinal.cpp" "$TMP/ref_linear.cpp" \
    "$TMP/ref_diffusive.cpp" "$TMP/ref_matrix.cpp" "$TMP/ref_matrix_secant.cpp"
echo "build_ref: built $OUT/libtroute_ref.so"
HIPCC=${HIPCC:-/opt/rocm/bin/hipcc}
"$HIPCC" -x hip --offload-host-only -O2 -ffp-contract=off -fPIC -shared -std=c++17 -w -Wl,--as-needed \
    -o "$OUT/libtroute_ref_f32.so" "$TMP/ref_wavefront_f32.cpp"
echo "build_ref: built $OUT/libtroute_ref_f32.so"
# the forcing-ingest program compiles whole from its own file (std C++17 only)
QLAT="$REF/wavefront/data_parsers/parse_qlat_fast_cpp.cpp"
if [ -f "$QLAT" ]; then
    g++ -O2 -std=c++17 -pthread -w -o "$OUT/parse_qlat_fast" "$QLAT"
    echo "build_ref: built $OUT/parse_qlat_fast"
fi
