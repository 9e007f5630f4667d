#!/bin/bash
# Regression run of program_gpu on GPU-BSW's own regression inputs.
#
#   regression_test.sh <GPU-BSW checkout> [work dir]
#
# Uses <checkout>/test-data/{dna,protein}-{reference,query}.fasta and compares the
# first column (the score) of program_gpu's output with
# <checkout>/regression_test/test_regression_{dna,protein}_good_result, which hold
# one score per pair.  (The checkout's quick_regression_test.sh diffs whole lines,
# but its sample program writes five columns per pair, so that diff cannot pass.)
# Exit status 0 when every score matches.
set -u
src=${1:?usage: regression_test.sh <GPU-BSW checkout> [work dir]}
work=${2:-.}
here=$(cd "$(dirname "$0")" && pwd)
prog="$here/program_gpu"
[ -x "$prog" ] || { echo "program_gpu is not built (make -C $here)"; exit 2; }
status=0
for kind in dna protein; do
    mode=$kind; [ $kind = protein ] && mode=aa
    ref="$src/test-data/$kind-reference.fasta"
    qry="$src/test-data/$kind-query.fasta"
    good="$src/regression_test/test_regression_${kind}_good_result"
    if [ ! -f "$ref" ] || [ ! -f "$qry" ] || [ ! -f "$good" ]; then
        echo "$kind: inputs missing ($ref, $qry, $good)"; status=2; continue
    fi
    out="$work/test_regression_${kind}_out"
    "$prog" $mode "$ref" "$qry" "$out" > "$work/test_regression_${kind}.log" 2>&1 || { echo "$kind: program_gpu failed"; status=1; continue; }
    if cut -f1 "$out" | diff -q - "$good" > /dev/null; then
        echo "$kind: $(wc -l < "$good") scores match"
    else
        echo "$kind: scores differ:"; cut -f1 "$out" | diff - "$good" | head -20; status=1
    fi
done
[ $status -eq 0 ] && echo "REGRESSION TESTS PASS"
exit $status
