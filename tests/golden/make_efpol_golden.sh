#!/bin/bash
# Writes the efpol golden tuv fixtures under tests/golden/efpol/ by running the
# reference efpol binaries built from the reference sources by
# oracle/Makefile.ref (outputs in oracle/_ref/).  Run from the repo root, in a
# container where the reference sources are present:
#     make -f oracle/Makefile.ref && bash tests/golden/make_efpol_golden.sh
# Each case directory keeps the reference argv in ARGV and the files the
# reference wrote (t=*_dev=0.res, simulation_settings.txt).
set -euo pipefail
ROOT="$(cd "$(dirname "$0")/../.." && pwd)"
OUT="$ROOT/tests/golden/efpol"
REF="$ROOT/oracle/_ref"
TMP="$(mktemp -d)"
trap 'rm -rf "$TMP"' EXIT
rm -rf "$OUT"
mkdir -p "$OUT"

# name  exec  SEED DENSITY TIME INTERVAL NMONO L MODEL
cases=(
  "ring_L16_N32_s1_equal   efpol_ring 1 1.2 300 100 32 16 sl_equal"
  "ring_L16_N32_s2_equal   efpol_ring 2 1.2 300 100 32 16 sl_equal"
  "ring_L16_N32_s3_double  efpol_ring 3 1.2 200 100 32 16 sl_double"
  "ring_L16_N32_s4_quad    efpol_ring 4 1.2 200 100 32 16 sl_quad"
  "ring_L32_N100_s1_equal  efpol_ring 1 1.2 100 100 100 32 sl_equal"
  "lin_L16_N32_s1_equal    efpol_lin  1 1.2 300 100 32 16 sl_equal"
  "lin_L16_N24_s5_quad     efpol_lin  5 1.0 200 100 24 16 sl_quad"
)
for c in "${cases[@]}"; do
  read -r name exe seed dens time intv nmono L model <<<"$c"
  d="$TMP/$name"
  (cd "$TMP" && "$REF/$exe" "$seed" "$d" "$dens" "$time" "$intv" "$nmono" 0 "$L" "$model" >/dev/null)
  mkdir -p "$OUT/$name"
  cp "$d/N_$nmono/"t=*_dev=0.res "$d/N_$nmono/simulation_settings.txt" "$OUT/$name/"
  echo "$exe $seed DIR $dens $time $intv $nmono 0 $L $model" > "$OUT/$name/ARGV"
done
du -sh "$OUT"
