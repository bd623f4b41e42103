#!/bin/bash
# Profiles of the bench workload (uneven 10k-cell batch) for the committed summaries:
#   1. rocprofv3 kernel trace + per-kernel stats of `bench.py --roofline-only --no-cpu`
#   2. HBM bytes read per dispatch (FETCH_SIZE) of the same command
#   3. HBM bytes written per dispatch (WRITE_SIZE) of the same command
# Raw output goes to $OUT (default profile_out/rocprof under the repository root).  Run from the
# repository root on a machine with the GPU; summarise with python_amd/profile_summary.py.
set -e
R=$(cd "$(dirname "$0")/.." && pwd)
OUT=${OUT:-$R/profile_out/rocprof}
mkdir -p "$OUT"
cd /tmp
export TMPDIR=/tmp
timeout -k 10 300 rocprofv3 --kernel-trace --stats --output-format csv -d "$OUT/trace" -o bench \
  -- python3 "$R/bench.py" --roofline-only --no-cpu > "$OUT/trace.log" 2>&1
echo "[profile] kernel trace done"
timeout -s KILL 150 rocprofv3 --pmc FETCH_SIZE --output-format csv -d "$OUT/fetch" -o bench \
  -- python3 "$R/bench.py" --roofline-only --no-cpu > "$OUT/fetch.log" 2>&1
echo "[profile] FETCH_SIZE done"
timeout -s KILL 150 rocprofv3 --pmc WRITE_SIZE --output-format csv -d "$OUT/write" -o bench \
  -- python3 "$R/bench.py" --roofline-only --no-cpu > "$OUT/write.log" 2>&1
echo "[profile] WRITE_SIZE done"
