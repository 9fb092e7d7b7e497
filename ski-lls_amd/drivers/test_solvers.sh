#!/bin/bash
# Runs the drivers over the reference's fixtures and prints the residuals in
# the format of the reference's test/test_solvers.sh ("<solver> gives a
# solution with residual <v>"), with the expected residuals of the
# rank-deficient set (test/test_solvers.sh:28-29).
HERE=$(cd "$(dirname "$0")" && pwd)
DATA=${1:-$HERE/../../tests/golden/data}
BIN=$HERE/bin

dense=(50by15DenseFullRank.mtx 50by20DenseFullRank.mtx 100by15DenseFullRank.mtx
       100by20DenseFullRank.mtx 120by15DenseFullRank.mtx 120by20DenseFullRank.mtx)
sparse=(50by15SparseFullRank.mtx 50by20SparseFullRank.mtx 100by15SparseFullRank.mtx
        100by20SparseFullRank.mtx 120by15SparseFullRank.mtx 120by20SparseFullRank.mtx)
rankdef=(lp_ship12l Franz1 GL7d26 cis-n4c6-b2 lp_modszk1 rel5 ch5-5-b1 n3c5-b2 ch4-4-b1
         n3c5-b1 n3c4-b1 cis-n4c6-b3)
expect=(18.3360 26.5028 50.8747 6.12592e-14 33.2357 14.0199 7.31938 8.99553e-15 4.23281
        3.46410 1.82574 30.9959)

run() { # solver matrix
    res=$("$BIN/$1" "$2" | grep Residual | awk '{print $NF}')
    echo "$1 gives a solution with residual $res"
}

for f in "${dense[@]}"; do echo "TESTING $f"; run ls_blendenpik_hashing "$DATA/$f"; echo; done
for f in "${sparse[@]}"; do
    echo "TESTING $f"
    run ls_qr "$DATA/$f"
    run ls_blendenpik_hashing_sparse_input "$DATA/$f"
    echo
done
for i in "${!rankdef[@]}"; do
    f=${rankdef[$i]}
    echo "TESTING $f.mtx"
    echo "We should get ${expect[$i]}"
    run ls_blendenpik_hashing_sparse_input "$DATA/rank_deficient/$f.mtx"
    run ls_qr "$DATA/rank_deficient/$f.mtx"
    echo
done
