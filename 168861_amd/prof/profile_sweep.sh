#!/bin/bash
# rocprofv3 runs behind the profiles of the 4,096-tree batched sweep (bench.py's
# sweep leg: FP64 tree solve, FP32 tree solve with FP64 refinement, explicit):
#   1. kernel trace + stats;
#   2. FETCH_SIZE, then WRITE_SIZE (one counter pass per run);
#   3. the same two counters on copy kernels of known size (8- and 16-byte accesses),
#      the calibration of both;
# then summarize.py reduces the CSV output to a small JSON (argument 1). Raw output
# goes under $PROF_RAW (default /tmp/nrt_prof). Run from the repository root, with
# the library and prof/copy_calib built.
set -o pipefail
export TMPDIR=/tmp
OUT=${1:?usage: profile_sweep.sh OUT.json [STEPS]}
STEPS=${2:-40}
RAW=${PROF_RAW:-/tmp/nrt_prof}
HERE=$(cd "$(dirname "$0")" && pwd)
mkdir -p "$RAW"
B="python3 bench.py --legs sweep --steps $STEPS --warmup 5 --no-oracle --no-cpu-baseline"
echo "[prof] kernel trace: $B"
timeout -k 10 600 rocprofv3 --kernel-trace --stats --output-format csv -d "$RAW/trace" -o sweep -- $B > "$RAW/bench_trace.json" 2> "$RAW/trace.err" || exit $?
echo "[prof] FETCH_SIZE"
timeout -s KILL 900 rocprofv3 --pmc FETCH_SIZE --output-format csv -d "$RAW/fetch" -o sweep -- $B > /dev/null 2> "$RAW/fetch.err" || exit $?
echo "[prof] WRITE_SIZE"
timeout -s KILL 900 rocprofv3 --pmc WRITE_SIZE --output-format csv -d "$RAW/write" -o sweep -- $B > /dev/null 2> "$RAW/write.err" || exit $?
echo "[prof] copy calibration"
timeout -s KILL 120 rocprofv3 --pmc FETCH_SIZE --output-format csv -d "$RAW/copy_fetch" -o copy -- "$HERE/copy_calib" > "$RAW/copy.txt" 2>&1 || exit $?
timeout -s KILL 120 rocprofv3 --pmc WRITE_SIZE --output-format csv -d "$RAW/copy_write" -o copy -- "$HERE/copy_calib" > /dev/null 2>&1 || exit $?
python3 "$HERE/summarize.py" "$RAW" "$OUT" --timed-steps "$STEPS" || exit $?
echo "[prof] done"
