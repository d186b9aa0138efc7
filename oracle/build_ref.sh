#!/usr/bin/env bash
# Builds the reference Fortran solver (pandaPQY/Moving-Mesh-reconstruction, mmh/) from its own
# sources, unmodified, into shared libraries under oracle/_ref/ for use as a parity checker and
# as the CPU baseline.  The reference fixes its grid size NG at compile time in the one-line
# configuration header dimen.f90 ("#define NG 512"); its workflow is to edit that header per run.
# Here each build directory holds symlinks to the reference sources plus a generated dimen.f90 that
# sets NG, so the unmodified sources compile for each NG.
#
#   usage: oracle/build_ref.sh [NG ...]      (default: 16 32 64)
# Outputs (git-ignored): oracle/_ref/libmmhref_ng<NG>_f32.so   FP32 as shipped (-O2, serial)
#                        oracle/_ref/libmmhref_ng<NG>_f64.so   -fdefault-real-8 promotion
#                        oracle/_ref/libmmhref_ng<NG>_f32omp.so FP32, -O3 -fopenmp (CPU baseline)
#                        oracle/_ref/dropin_ref_ng<NG>        tests/dropin/dropin_driver.F90 linked
#                                                             against libmmhref_ng<NG>_f32.so
#                        oracle/_ref/libmmhref_fisher.so      calcfisherinfo.f90 (its calccov,
#                        calerror and inverse subroutines are called; the program itself is not run)
set -euo pipefail
REF=${MMH_REFERENCE_DIR:-/root/reference}/mmh
HERE=$(cd "$(dirname "$0")" && pwd)
OUT=$HERE/_ref
FC=${FC:-flang}
if [ ! -d "$REF" ]; then echo "build_ref: $REF not present; skipping"; exit 0; fi
command -v "$FC" >/dev/null || { echo "build_ref: $FC not found; skipping"; exit 0; }
NGS=${*:-16 32 64}
SRCS="gauss1 mgutil multigrid limiter relaxing recipes12 gradient reconstruct"
INCS="dimen relaxgl globalpa cold mg8 mg16 mg32 mg64 mg128 mg256 mg512 mg1024"
mkdir -p "$OUT/build"
for NG in $NGS; do
  # REF_VARIANTS selects the builds (the CPU baseline at NG = 512 needs only f32omp)
  for V in ${REF_VARIANTS:-f32 f64 f32omp}; do
    B=$OUT/build/ng${NG}_$V
    rm -rf "$B"; mkdir -p "$B"
    for f in $SRCS $INCS; do ln -s "$REF/$f.f90" "$B/$f.f90"; done
    rm "$B/dimen.f90"; echo "#define NG $NG" > "$B/dimen.f90"
    case $V in
      f32) FL="-O2";; f64) FL="-O2 -fdefault-real-8";; f32omp) FL="-O3 -fopenmp";;
    esac
    # the reference's own Makefile builds with -mcmodel=large (mmh/Makefile:7); it is needed from
    # NG = 512 up, where the static work arrays pass 2 GiB
    [ "$NG" -ge 512 ] && FL="$FL -mcmodel=large"
    objs=""
    for f in $SRCS; do
      # gradient/reconstruct declare explicit real(4) arrays and cannot be promoted consistently
      if [ $V = f64 ] && { [ $f = gradient ] || [ $f = reconstruct ]; }; then continue; fi
      (cd "$B" && $FC -cpp -DCOLD -DOMPTIME -fPIC $FL -w -c "$f.f90" -o "$f.o")
      objs="$objs $B/$f.o"
    done
    $FC -shared -fPIC $FL $objs -o "$OUT/libmmhref_ng${NG}_$V.so"
    echo "built $OUT/libmmhref_ng${NG}_$V.so"
  done
  # the reference-style caller of tests/dropin/, linked against this build's objects (FP32 as shipped)
  [ -f "$OUT/libmmhref_ng${NG}_f32.so" ] || continue
  $FC -cpp -DNG=$NG -O2 -w "$HERE/../tests/dropin/dropin_driver.F90" -o "$OUT/dropin_ref_ng$NG" \
    -L "$OUT" -lmmhref_ng${NG}_f32 -Wl,-rpath,'$ORIGIN'
  echo "built $OUT/dropin_ref_ng$NG"
done
# the Fisher-matrix statistics (independent of NG)
B=$OUT/build/fisher
rm -rf "$B"; mkdir -p "$B"
ln -s "$REF/calcfisherinfo.f90" "$B/calcfisherinfo.f90"
(cd "$B" && $FC -O2 -fPIC -w -c calcfisherinfo.f90 -o calcfisherinfo.o)
$FC -shared -fPIC "$B/calcfisherinfo.o" -o "$OUT/libmmhref_fisher.so"
echo "built $OUT/libmmhref_fisher.so"
