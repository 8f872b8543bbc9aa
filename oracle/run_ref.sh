#!/bin/bash
# Runs one compiled reference CPU program (built by `make -C oracle ref`) at its
# hard-coded configuration.  The programs read ../../../data/<file> and write
# ../../../results/<name>, relative to their working directory, and wait on
# stdin at the end; this script gives them a scratch tree under oracle/_ref/run
# whose data/ links to the reference's sample catalogs and whose results/
# collects the outputs.
#   usage: oracle/run_ref.sh <prog> <arg1> [<arg2> [<arg3>]]
# Outputs land in oracle/_ref/run/results/.
set -euo pipefail
HERE="$(cd "$(dirname "$0")" && pwd)"
REF="${REF:-/root/reference}"
prog="$1"; shift
run="$HERE/_ref/run"
mkdir -p "$run/a/b/c" "$run/results"
ln -sfn "$REF/data" "$run/data"
cd "$run/a/b/c"
"$HERE/_ref/$prog.out" "$@" </dev/null >"$run/$prog.log" 2>&1
