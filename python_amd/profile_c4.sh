#!/bin/bash
# Profiles of the macro-atom inverse (C4: 2,000 cells N = I - Q, n in {256, 512, 1024}):
#   1. rocprofv3 kernel trace + per-kernel stats of `bench.py --config c4 --steps 1` (both plans)
#   2. MFMA counters per dispatch of the FP64 plan on a tenth of the batch (--quick):
#      SQ_VALU_MFMA_BUSY_CYCLES, GRBM_GUI_ACTIVE, SQ_INSTS_VALU_MFMA_MOPS_F64, _MOPS_BF16
# Raw output goes to $OUT (default profile_out/prof_c4 under the repository root); summarise with
# python -m python_amd.profile_summary c4 [OUT].
set -e
R=$(cd "$(dirname "$0")/.." && pwd)
OUT=${OUT:-$R/profile_out/prof_c4}
mkdir -p "$OUT"
cd /tmp
export TMPDIR=/tmp
timeout -k 10 400 rocprofv3 --kernel-trace --stats --output-format csv -d "$OUT/trace" -o c4 \
  -- python3 "$R/bench.py" --config c4 --steps 1 > "$OUT/trace.log" 2>&1
echo "[profile c4] kernel trace done"
timeout -s KILL 300 rocprofv3 --pmc SQ_VALU_MFMA_BUSY_CYCLES GRBM_GUI_ACTIVE SQ_INSTS_VALU_MFMA_MOPS_F64 \
  SQ_INSTS_VALU_MFMA_MOPS_BF16 --output-format csv -d "$OUT/mfma" -o c4 \
  -- python3 "$R/bench.py" --config c4 --steps 1 --quick > "$OUT/mfma.log" 2>&1
echo "[profile c4] MFMA counters done"
