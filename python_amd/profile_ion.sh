#!/bin/bash
# Kernel trace of the lock-step ionization iteration (10k uneven cells, seed 6): rocprofv3
# --kernel-trace --stats of `bench.py --config ion --steps 1 --warmup 1 --no-cpu` (three whole
# populations() calls: warm-up, timed, checked).  Raw output goes to $OUT (default
# profile_out/ion under the repository root); summarise with
# python -m python_amd.profile_summary ion [OUT].
set -e
R=$(cd "$(dirname "$0")/.." && pwd)
OUT=${OUT:-$R/profile_out/ion}
mkdir -p "$OUT"
cd /tmp
export TMPDIR=/tmp
timeout -k 10 400 rocprofv3 --kernel-trace --stats --output-format csv -d "$OUT/trace" -o ion \
  -- python3 "$R/bench.py" --config ion --steps 1 --warmup 1 --no-cpu > "$OUT/trace.log" 2>&1
echo "[profile ion] kernel trace done"
