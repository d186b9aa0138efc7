#!/usr/bin/env bash
# Links tests/dropin/dropin_driver.F90 against the GPU drop-ins (libmmh_fortran.so, built by
# moving-mesh-reconstruction_amd/build.py) into tests/dropin/_bin/dropin_gpu_ng<NG>.  The copy linked
# against the reference's own objects is built by oracle/build_ref.sh.   usage: build.sh [NG ...]
set -euo pipefail
HERE=$(cd "$(dirname "$0")" && pwd)
PKG=$HERE/../../moving-mesh-reconstruction_amd
FC=${FC:-flang}
command -v "$FC" >/dev/null || { echo "dropin: $FC not found; skipping"; exit 0; }
[ -f "$PKG/libmmh_fortran.so" ] || { echo "dropin: libmmh_fortran.so not built; skipping"; exit 0; }
mkdir -p "$HERE/_bin"
for NG in ${*:-32}; do
  "$FC" -cpp -DNG=$NG -DMMH_GPU -O2 -I "$PKG/_build" "$HERE/dropin_driver.F90" -o "$HERE/_bin/dropin_gpu_ng$NG" \
    -L "$PKG" -lmmh_fortran -lmmh_amd -Wl,-rpath,'$ORIGIN/../../../moving-mesh-reconstruction_amd'
  echo "built $HERE/_bin/dropin_gpu_ng$NG"
done
