#!/bin/bash
# Regenerates tests/golden/ref_out/<program>/ from the reference CPU programs
# compiled unmodified by `make -C oracle ref` (outputs of the reference's own
# code on the reference's sample catalogs, at the programs' hard-coded
# configurations).  Needs /root/reference; run in this container only.
set -euo pipefail
ROOT="$(cd "$(dirname "$0")/../.." && pwd)"
OUT="$ROOT/tests/golden/ref_out"
RES="$ROOT/oracle/_ref/run/results"
make -C "$ROOT/oracle" ref >/dev/null
run() {
  local prog="$1"; shift
  rm -rf "$RES"; mkdir -p "$RES"
  "$ROOT/oracle/run_ref.sh" "$prog" "$@"
  mkdir -p "$OUT/$prog"
  cp "$RES"/* "$OUT/$prog/"
  echo "$prog $*" > "$OUT/$prog/COMMAND"
}
run 2iso_normal data.dat rand0.dat ref
run 2iso_bpc data.dat rand0.dat ref
run 2iso_bpca data.dat ref
run 2ani_normal data.dat rand0.dat ref
run 2ani_bpc data.dat rand0.dat ref
run 3iso_normal data_10K.dat rand0_10K.dat ref
# cpp/3PCFisotropic/BPC computes only ff_av / ff_av_ref of the analytic path
# (its RRR/DDR loop is commented out) and writes them to its working directory.
CWD="$ROOT/oracle/_ref/run/a/b/c"
rm -f "$CWD"/ff_av.dat "$CWD"/ff_av_ref.dat "$CWD"/DD.dat
"$ROOT/oracle/run_ref.sh" 3iso_bpc data.dat ref
mkdir -p "$OUT/3iso_bpca"
cp "$CWD/ff_av.dat" "$CWD/ff_av_ref.dat" "$OUT/3iso_bpca/"
echo "3iso_bpc data.dat ref (writes ff_av.dat, ff_av_ref.dat to its working directory)" \
  > "$OUT/3iso_bpca/COMMAND"
# cpp/3PCFanisotropic/normal writes ~50 MB histograms to its working directory;
# its outputs are not kept as fixtures.
