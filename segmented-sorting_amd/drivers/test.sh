#!/bin/bash
# Golden diff of a driver over a directory of inputs, in the manner of the
# reference's scripts/test.sh: for every <name>.in in IN_DIR, run
# `PROG [ARGS...] < IN_DIR/<name>.in` and compare its stdout byte for byte with
# EXPECT_DIR/<name>.out; a mismatching output is kept as ERR_DIR/<name>.out.
# Prints one line per file and exits 1 if any file differs.
# usage: test.sh PROG IN_DIR EXPECT_DIR ERR_DIR [ARGS...]   (e.g. ARGS = --values)
prog=$1; in_dir=$2; exp_dir=$3; err_dir=$4
shift 4
bad=0
for f in "$in_dir"/*.in; do
    name=$(basename "$f" .in)
    out=$(mktemp)
    if ! "$prog" "$@" < "$f" > "$out"; then
        echo "$name: $prog failed"
        bad=1
    elif cmp -s "$out" "$exp_dir/$name.out"; then
        echo "$name: ok"
    else
        mkdir -p "$err_dir"
        cp "$out" "$err_dir/$name.out"
        echo "$name: differs (kept in $err_dir/$name.out)"
        bad=1
    fi
    rm -f "$out"
done
exit $bad
