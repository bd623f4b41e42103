#!/bin/bash
# Kernel timeline of the bench workload's headline plan (uneven 10k-cell batch, mixed mode,
# autotuned, graph replays): rocprofv3 --kernel-trace of
# `python3 -m python_amd.profile_summary run_timeline` (three warm-up solves, a spin-kernel marker,
# two marked solves).  Raw output goes to $OUT (default profile_out/timeline under the repository
# root); summarise with python -m python_amd.profile_summary timeline [OUT].
set -e
R=$(cd "$(dirname "$0")/.." && pwd)
OUT=${OUT:-$R/profile_out/timeline}
mkdir -p "$OUT"
cd /tmp
export TMPDIR=/tmp
export PYTHONPATH="$R${PYTHONPATH:+:$PYTHONPATH}"
timeout -k 10 300 rocprofv3 --kernel-trace --output-format csv -d "$OUT/trace" -o tl \
  -- python3 -m python_amd.profile_summary run_timeline > "$OUT/trace.log" 2>&1
echo "[profile timeline] kernel trace done"
