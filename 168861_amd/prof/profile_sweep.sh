#!/bin/bash
# rocprofv3 runs behind the profiles of the 4,096-tree batched sweep (bench.py's
# sweep leg: FP64 tree solve, FP32 tree solve with FP64 refinement, explicit).
# Usage: profile_sweep.sh OUT.json STEPS PHASE...  with PHASE one of
#   trace    kernel trace + stats;
#   fetch    FETCH_SIZE (one counter pass per run);
#   write    WRITE_SIZE;
#   dram     TCC_EA0_RDREQ_DRAM_32B_sum and TCC_EA0_WRREQ_WRITE_DRAM_32B_sum (HBM-side
#            read and write requests in 32-byte units: a cross-check of the two above);
#   sq       wave-state counters (SQ_WAVE_CYCLES, SQ_WAIT_ANY, SQ_WAIT_INST_ANY,
#            SQ_ACTIVE_INST_ANY, SQ_ACTIVE_INST_VALU, ...): where the waves' cycles go;
#   calib    the three counter runs on copy kernels of known size (8- and 16-byte lanes);
#   summary  summarize.py reduce: what this box's runs left, reduced to OUT.json
#            (runs on separate boxes are combined by summarize.py merge).
# Raw output goes under $PROF_RAW (default /tmp/nrt_prof); the bench's progress
# lines are appended to $PROF_LOG. Run from the repository root, with the library
# and prof/copy_calib built.
set -o pipefail
export TMPDIR=/tmp
OUT=${1:?usage: profile_sweep.sh OUT.json STEPS PHASE...}
STEPS=${2:?steps}
shift 2
RAW=${PROF_RAW:-/tmp/nrt_prof}
HERE=$(cd "$(dirname "$0")" && pwd)
mkdir -p "$RAW"
# The trace run tunes the launch shapes and keeps them in $SHAPES (unless that
# file exists already); the counter runs reuse them, so every run times the same
# kernels and the counter runs skip the tuning.
SHAPES=${SHAPES:-$RAW/shapes.json}
LOG=${PROF_LOG:-/dev/null}
DRAM="TCC_EA0_RDREQ_DRAM_32B_sum TCC_EA0_WRREQ_WRITE_DRAM_32B_sum"
SQ="SQ_WAVES SQ_WAVE_CYCLES SQ_BUSY_CYCLES SQ_WAIT_ANY SQ_WAIT_INST_ANY SQ_ACTIVE_INST_ANY SQ_ACTIVE_INST_VALU SQ_INSTS_VALU"
B="python3 bench.py --legs sweep --steps $STEPS --warmup 5 --no-oracle --no-cpu-baseline --shapes $SHAPES"
for ph in "$@"; do
  case $ph in
    trace)
      echo "[prof] kernel trace: $B"
      timeout -k 10 600 rocprofv3 --kernel-trace --stats --output-format csv -d "$RAW/trace" -o sweep -- $B > "$RAW/bench_trace.json" 2>> "$LOG" || exit $?
      ;;
    fetch|write)
      C=$([ "$ph" = fetch ] && echo FETCH_SIZE || echo WRITE_SIZE)
      echo "[prof] $C: $B"
      timeout -s KILL 900 rocprofv3 --pmc $C --output-format csv -d "$RAW/$ph" -o sweep -- $B > /dev/null 2>> "$LOG" || exit $?
      ;;
    dram)
      echo "[prof] DRAM request counters: $B"
      timeout -s KILL 900 rocprofv3 --pmc $DRAM --output-format csv -d "$RAW/dram" -o sweep -- $B > /dev/null 2>> "$LOG" || exit $?
      ;;
    sq)
      echo "[prof] wave-state counters: $B"
      timeout -s KILL 900 rocprofv3 --pmc $SQ --output-format csv -d "$RAW/sq" -o sweep -- $B > /dev/null 2>> "$LOG" || exit $?
      ;;
    calib)
      echo "[prof] copy calibration"
      timeout -s KILL 120 rocprofv3 --pmc FETCH_SIZE --output-format csv -d "$RAW/copy_fetch" -o copy -- "$HERE/copy_calib" > "$RAW/copy.txt" 2>&1 || exit $?
      timeout -s KILL 120 rocprofv3 --pmc WRITE_SIZE --output-format csv -d "$RAW/copy_write" -o copy -- "$HERE/copy_calib" >> "$RAW/copy.txt" 2>&1 || exit $?
      timeout -s KILL 120 rocprofv3 --pmc $DRAM --output-format csv -d "$RAW/copy_dram" -o copy -- "$HERE/copy_calib" >> "$RAW/copy.txt" 2>&1 || exit $?
      ;;
    summary)
      python3 "$HERE/summarize.py" reduce "$RAW" "$OUT" --timed-steps "$STEPS" || exit $?
      ;;
    *) echo "unknown phase $ph"; exit 2 ;;
  esac
done
echo "[prof] done: $*"
