#!/bin/bash
# Stall counters of the bench workload's kernels (uneven 10k-cell batch, every FP64 size class
# launched in isolation by `bench.py --roofline-only --no-cpu`), in two counter passes:
#   1. SQ_WAVE_CYCLES SQ_WAIT_ANY SQ_WAIT_INST_ANY SQ_ACTIVE_INST_VALU SQ_ACTIVE_INST_LDS
#      SQ_LDS_BANK_CONFLICT SQ_BUSY_CYCLES SQ_ACTIVE_INST_ANY
#   2. SQ_WAIT_INST_LDS SQ_INSTS_VALU SQ_INSTS_LDS SQ_LDS_ADDR_CONFLICT SQ_INST_CYCLES_VMEM_RD
#      SQ_INST_CYCLES_SALU SQ_WAVES GRBM_GUI_ACTIVE
# Raw output goes to $OUT (default profile_out/stalls under the repository root); summarise with
# python -m python_amd.profile_summary stalls [OUT].
set -e
R=$(cd "$(dirname "$0")/.." && pwd)
OUT=${OUT:-$R/profile_out/stalls}
mkdir -p "$OUT"
cd /tmp
export TMPDIR=/tmp
timeout -s KILL 200 rocprofv3 --pmc SQ_WAVE_CYCLES SQ_WAIT_ANY SQ_WAIT_INST_ANY SQ_ACTIVE_INST_VALU SQ_ACTIVE_INST_LDS \
  SQ_LDS_BANK_CONFLICT SQ_BUSY_CYCLES SQ_ACTIVE_INST_ANY --output-format csv -d "$OUT/p1" -o st \
  -- python3 "$R/bench.py" --roofline-only --no-cpu > "$OUT/p1.log" 2>&1
echo "[profile stalls] pass 1 done"
timeout -s KILL 200 rocprofv3 --pmc SQ_WAIT_INST_LDS SQ_INSTS_VALU SQ_INSTS_LDS SQ_LDS_ADDR_CONFLICT \
  SQ_INST_CYCLES_VMEM_RD SQ_INST_CYCLES_SALU SQ_WAVES GRBM_GUI_ACTIVE --output-format csv -d "$OUT/p2" -o st \
  -- python3 "$R/bench.py" --roofline-only --no-cpu > "$OUT/p2.log" 2>&1
echo "[profile stalls] pass 2 done"
