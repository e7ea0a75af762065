#!/usr/bin/env bash
# Builds the reference Fortran solver, read in place from the reference
# checkout, once per case header in oracle/ref_cases/, together with
# oracle/ref_harness.f90, into oracle/_ref/<case>/ref_harness.
#
# TEST INFRASTRUCTURE ONLY (the checker and the CPU baseline).  The reference
# selects its options with compile-time macros from ../../config.h relative to
# each source file, so every case gets a scratch tree whose source/ files are
# symlinks to the reference's files and whose config.h is the case header:
# nothing of the reference is copied.  Only the executables land in
# oracle/_ref/ (git-ignored).  Toolchain: flang (ROCm) + MKL from /opt/conda.
set -euo pipefail

REF=${LSSPH_REFERENCE:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
OUT=$HERE/_ref
FC=${FC:-flang}
MKL=${MKL_LIB:-/opt/conda/lib}

if [ ! -d "$REF/source" ]; then
  echo "build_ref.sh: reference not found at $REF; skipping" >&2
  exit 0
fi

# Module order of the reference Makefile (Makefile:70-100), without main.f90.
SRCS=(
  source/core/global_types.f90
  source/core/kernel_functions_mod.f90
  source/io/file_operations_mod.f90
  source/check/check_param_mod.f90
  source/setup/init_param_mod.f90
  source/setup/setup_closed_box_mod.f90
  source/setup/impose_initial_conditions_mod.f90
  source/setup/restart_mod.f90
  source/setup/init_particle_mod.f90
  source/setup/init_virtual_markers_mod.f90
  source/neighbor/init_cell_mod.f90
  source/neighbor/update_cell_mod.f90
  source/boundary/interpolate_CSPH_mod.f90
  source/boundary/interpolate_LSSPH_B_mod.f90
  source/boundary/extrapolate_VM_to_WL_mod.f90
  source/boundary/extrapolate_diff_test_mod.f90
  source/boundary/ghost_mod.f90
  source/equation/classical_SPH_mod.f90
  source/equation/LSSPH_A_mod.f90
  source/equation/calculate_RHS_mod.f90
  source/equation/RK2_mod.f90
  source/equation/RK4_mod.f90
  source/shifting/shift_interpolate_mod.f90
  source/shifting/PST_mod.f90
  source/check/check_steady_state_mod.f90
)

CASES=("$@")
if [ ${#CASES[@]} -eq 0 ]; then
  for f in "$HERE"/ref_cases/*.h; do
    n=$(basename "$f" .h)
    [ "$n" = base ] || CASES+=("$n")
  done
fi

build_case() {
  local name=$1
  local hdr=$HERE/ref_cases/$name.h
  local flags
  flags=$(sed -n 's#^/\* FFLAGS: \(.*\) \*/$#\1#p' "$hdr")
  flags=${flags:-"-O2"}
  local skip
  skip=$(sed -n 's#^/\* SKIP: \(.*\) \*/$#\1#p' "$hdr")
  local tmp
  tmp=$(mktemp -d /tmp/lssph_ref_XXXXXX)
  cp "$hdr" "$tmp/config.h"
  cp "$HERE/ref_cases/base.h" "$tmp/base.h"
  mkdir -p "$tmp/build"
  local objs=()
  for s in "${SRCS[@]}"; do
    case " $skip " in *" $s "*) continue ;; esac
    mkdir -p "$tmp/$(dirname "$s")"
    ln -s "$REF/$s" "$tmp/$s"
    local o="$tmp/build/$(basename "$s" .f90).o"
    (cd "$tmp" && $FC $flags -fopenmp -cpp -module-dir build -I build -I . -c "$s" -o "$o")
    objs+=("$o")
  done
  mkdir -p "$OUT/$name"
  (cd "$tmp" && $FC $flags -fopenmp -cpp -module-dir build -I build -I . \
      "$HERE/ref_harness.f90" "${objs[@]}" -o "$OUT/$name/ref_harness" \
      -L"$MKL" -Wl,-rpath,"$MKL" -lmkl_intel_lp64 -lmkl_sequential -lmkl_core \
      -lpthread -lm -ldl)
  rm -rf "$tmp"
  echo "built $OUT/$name/ref_harness ($flags)"
}

pids=()
for c in "${CASES[@]}"; do
  build_case "$c" &
  pids+=($!)
done
rc=0
for p in "${pids[@]}"; do
  wait "$p" || rc=1
done
exit $rc
