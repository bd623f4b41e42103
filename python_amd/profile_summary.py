"""Summarise the rocprofv3 output of python_amd/profile_bench.sh into profiles/.

    python -m python_amd.profile_summary [OUT_DIR]      (default profile_out/rocprof)
    python -m python_amd.profile_summary c4 [OUT_DIR]   (python_amd/profile_c4.sh output)
    python -m python_amd.profile_summary stalls [OUT_DIR] (python_amd/profile_stalls.sh output)
    python -m python_amd.profile_summary ion [OUT_DIR]  (python_amd/profile_ion.sh output)
    python -m python_amd.profile_summary timeline [OUT_DIR] (python_amd/profile_timeline.sh output)

Reads the kernel statistics and kernel trace of `bench.py --roofline-only --no-cpu`, its
FETCH_SIZE and WRITE_SIZE counter runs and the roofline object the command printed, and writes
  profiles/<kernel>_uneven10k_trace.txt   per-kernel statistics, the dominant kernel's dispatches
                                          against the HIP-event average reported by bench.py
  profiles/<kernel>_uneven10k_hbm.txt     HBM bytes per dispatch, calibrated on a 512 MiB copy
  profiles/traffic.json                   the calibrated bytes per launch bench.py reports
"""
from __future__ import annotations

import csv
import json
import os
import re
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CMD = ("python_amd/profile_bench.sh  (rocprofv3 --kernel-trace --stats --output-format csv -- python3 bench.py "
       "--roofline-only --no-cpu; then the same command under --pmc FETCH_SIZE and under --pmc WRITE_SIZE)")


def _short(name: str) -> str:
    m = re.match(r"(?:void )?(\w+)<([^>]*)>", name)
    if not m:
        return re.sub(r"\W+", "_", name)[:40]
    return m.group(1) + "_" + re.sub(r"[^0-9a-z]+", "_", m.group(2).lower()).strip("_")


def _roofline(out):
    with open(os.path.join(out, "trace.log")) as f:
        for line in f:
            if line.startswith('{"roofline"'):
                return json.loads(line)
    raise RuntimeError("no roofline line in trace.log")


def _counters(path, counter):
    """kernel name -> list of per-dispatch values of one counter."""
    per = {}
    with open(path) as f:
        for r in csv.DictReader(f):
            if r["Counter_Name"] != counter:
                continue
            per.setdefault(r["Kernel_Name"], []).append(float(r["Counter_Value"]))
    return per


def _counter_seq(path, counter):
    """(kernel name, value) of one counter for every dispatch, in dispatch order."""
    rows = []
    with open(path) as f:
        for r in csv.DictReader(f):
            if r["Counter_Name"] == counter:
                rows.append((int(r["Dispatch_Id"]), r["Kernel_Name"], float(r["Counter_Value"])))
    rows.sort()
    return [(nm, v) for _, nm, v in rows]


def _trace_seq(trace):
    """(kernel name, duration ns) of every dispatch of a kernel trace, in start order."""
    rows = sorted(trace, key=lambda r: int(r["Start_Timestamp"]))
    return [(r["Kernel_Name"], int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) for r in rows]


def _unit_values(seq, parts, is_part):
    """Per part, the values of the dispatches that form the unit: a dispatch of the first part
    followed directly by one of each further part (a factor kernel and its back substitution),
    so that another unit launching the same factor kernel with a different back substitution is
    not counted.  A one-part unit takes every dispatch of its kernel."""
    out = [[] for _ in parts]
    i = 0
    while i < len(seq):
        if is_part(seq[i][0], parts[0]) and i + len(parts) <= len(seq) and all(
                is_part(seq[i + j][0], parts[j]) for j in range(1, len(parts))):
            for j in range(len(parts)):
                out[j].append(seq[i + j][1])
            i += len(parts)
        else:
            i += 1
    return out


def main(out=None):
    out = out or os.path.join(ROOT, "profile_out", "rocprof")
    rl = _roofline(out)
    roof = rl["roofline"]
    dom = roof["kernel"]
    calib_bytes = rl["calibration_copy_bytes"]

    stats = list(csv.DictReader(open(os.path.join(out, "trace", "bench_kernel_stats.csv"))))
    trace = list(csv.DictReader(open(os.path.join(out, "trace", "bench_kernel_trace.csv"))))
    # the measured unit can be a kernel pair ("factor + back substitution", launched back to back)
    parts = [x.strip() for x in dom.split(" + ")]

    def is_part(name, part):
        return name.startswith("void " + part + "(") or name.startswith(part + "(")

    tseq = _trace_seq(trace)
    per_part = _unit_values(tseq, parts, is_part)
    dom_d = per_part[0]
    os.makedirs(os.path.join(ROOT, "profiles"), exist_ok=True)
    short = _short(parts[0])

    lines = [f"# Kernel trace of the bench workload (uneven 10k-cell batch, FP64 size classes launched in isolation)",
             f"command: {CMD}", "",
             f"{'kernel':78s} {'calls':>6s} {'avg us':>10s} {'total ms':>10s} {'%':>6s}"]
    for r in stats[:14]:
        lines.append(f"{r['Name'][:78]:78s} {int(r['Calls']):6d} {float(r['AverageNs']) / 1e3:10.1f} "
                     f"{float(r['TotalDurationNs']) / 1e6:10.3f} {float(r['Percentage']):6.2f}")
    part_avg = [sum(d) / len(d) / 1e6 if d else float("nan") for d in per_part]
    avg_ms = sum(part_avg)
    lines += ["", f"dominant kernel: {dom}",
              f"  dispatches in the trace: {', '.join(f'{p}: {len(d)}' for p, d in zip(parts, per_part))} "
              f"(size classes {', '.join(roof['size_classes'])}, {roof['launches_per_step']} per repetition)",
              f"  rocprofv3 average duration: {' + '.join(f'{a:.4f}' for a in part_avg)} = {avg_ms:.4f} ms",
              f"  bench.py HIP-event average: {roof['kernel_avg_ms']:.4f} ms per launch",
              f"  algorithmic bytes per launch: {roof['algorithmic_bytes_per_launch']}, flops per launch: "
              f"{roof['algorithmic_flops_per_launch']}",
              f"  achieved: {roof['achieved']} {roof['unit']} of {roof['peak']} ({roof['bound']} bound), "
              f"frac {roof['frac']}; HBM {roof['hbm_achieved_gbs']} GB/s algorithmic"]
    with open(os.path.join(ROOT, "profiles", f"{short}_uneven10k_trace.txt"), "w") as f:
        f.write("\n".join(lines) + "\n")

    fetch = _counters(os.path.join(out, "fetch", "bench_counter_collection.csv"), "FETCH_SIZE")
    write = _counters(os.path.join(out, "write", "bench_counter_collection.csv"), "WRITE_SIZE")

    def pick(per, pat):
        for k, v in per.items():
            if pat(k):
                return k, v
        raise RuntimeError("kernel not found")

    cname, f_cal = pick(fetch, lambda k: "MulFunctor" in k)
    _, w_cal = pick(write, lambda k: "MulFunctor" in k)
    kib = calib_bytes / 1024.0
    ff = max(f_cal) / kib
    fw = max(w_cal) / kib
    fseq = _counter_seq(os.path.join(out, "fetch", "bench_counter_collection.csv"), "FETCH_SIZE")
    wseq = _counter_seq(os.path.join(out, "write", "bench_counter_collection.csv"), "WRITE_SIZE")
    f_kib = sum(sum(d) / len(d) for d in _unit_values(fseq, parts, is_part))
    w_kib = sum(sum(d) / len(d) for d in _unit_values(wseq, parts, is_part))
    f_dom = [f_kib]
    w_dom = [w_kib]
    fetch_b = f_kib * 1024.0 / ff
    write_b = w_kib * 1024.0 / fw
    traffic = fetch_b + write_b
    tj = {"kernel": dom, "hbm_bytes_per_launch": round(traffic), "fetch_bytes_per_launch": round(fetch_b),
          "write_bytes_per_launch": round(write_b), "algorithmic_bytes_per_launch": roof["algorithmic_bytes_per_launch"],
          "fetch_kib_per_launch_raw": round(f_kib, 1),
          "write_kib_per_launch_raw": round(w_kib, 1),
          "calibration": {"kernel": cname[:90], "bytes_read_and_written": calib_bytes,
                          "fetch_factor": round(ff, 4), "write_factor": round(fw, 4)},
          "command": CMD}
    with open(os.path.join(ROOT, "profiles", "traffic.json"), "w") as f:
        json.dump(tj, f, indent=1)
        f.write("\n")
    hl = [f"# HBM traffic of {dom} on the bench workload (per launch)",
          f"command: {CMD}", "",
          f"calibration: {cname[:90]} reads and writes {calib_bytes} B; FETCH_SIZE reads {max(f_cal):.0f} KiB "
          f"(factor {ff:.4f}), WRITE_SIZE {max(w_cal):.0f} KiB (factor {fw:.4f})",
          f"FETCH_SIZE {f_kib:.1f} KiB per launch -> {fetch_b / 1e6:.2f} MB",
          f"WRITE_SIZE {w_kib:.1f} KiB per launch -> {write_b / 1e6:.2f} MB",
          f"HBM traffic per launch: {traffic / 1e6:.2f} MB against {roof['algorithmic_bytes_per_launch'] / 1e6:.2f} MB "
          f"algorithmic ({traffic / roof['algorithmic_bytes_per_launch']:.2f}x)"]
    with open(os.path.join(ROOT, "profiles", f"{short}_uneven10k_hbm.txt"), "w") as f:
        f.write("\n".join(hl) + "\n")
    cl = _class_table(rl, tseq, fseq, wseq, ff, fw, is_part)
    with open(os.path.join(ROOT, "profiles", "size_classes_uneven10k.txt"), "w") as f:
        f.write("\n".join(cl) + "\n")
    print("\n".join(lines + [""] + hl + [""] + cl))


def _class_table(rl, tseq, fseq, wseq, ff, fw, is_part):
    """Every FP64 size class of the bench workload: rocprofv3 average duration of its kernels,
    algorithmic bytes, achieved GB/s and the calibrated FETCH_SIZE + WRITE_SIZE bytes."""
    out = ["# Size classes of the bench workload (uneven 10k-cell batch), each launched in isolation",
           f"command: {CMD}", "",
           "Per class: rocprofv3 average duration of each kernel of the class (factor, then back "
           "substitution where there is one), algorithmic bytes 8 (n^2 + 2 n) per cell, the achieved "
           "rate on those bytes, and the HBM bytes the counters saw (FETCH_SIZE / 0.5 + WRITE_SIZE, per "
           "dispatch).", "",
           f"{'kernel(s)':58s} {'cells':>6s} {'us':>9s} {'alg MB':>8s} {'GB/s':>8s} {'% HBM':>6s} "
           f"{'HBM MB':>8s} {'x alg':>6s}"]
    for c in rl.get("classes", []):
        parts = [x.strip() for x in c["kernel"].split(" + ")]
        # classes sharing a unit (the same kernels): the dispatches of this class's count
        same = [x for x in rl["classes"] if x["kernel"] == c["kernel"]]
        us = 0.0
        hbm = 0.0
        tv = _unit_values(tseq, parts, is_part)
        fvs = _unit_values(fseq, parts, is_part)
        wvs = _unit_values(wseq, parts, is_part)
        for j in range(len(parts)):
            d, fv, wv = tv[j], fvs[j], wvs[j]
            if len(same) > 1:
                k = same.index(c)
                d, fv, wv = d[k::len(same)], fv[k::len(same)], wv[k::len(same)]
            us += sum(d) / len(d) / 1e3 if d else float("nan")
            if fv and wv:
                hbm += (sum(fv) / len(fv)) * 1024.0 / ff + (sum(wv) / len(wv)) * 1024.0 / fw
        alg = c.get("bytes", 0.0)
        gbs = alg / (us * 1e-6) / 1e9 if us > 0 else float("nan")
        out.append(f"{c['kernel'][:58]:58s} {c['count']:6d} {us:9.1f} {alg / 1e6:8.2f} {gbs:8.1f} "
                   f"{100 * gbs / 8000:6.2f} {hbm / 1e6:8.2f} {hbm / alg if alg else float('nan'):6.2f}")
    return out


C4_CMD = ("python_amd/profile_c4.sh  (rocprofv3 --kernel-trace --stats --output-format csv -- python3 bench.py "
          "--config c4 --steps 1; then rocprofv3 --pmc SQ_VALU_MFMA_BUSY_CYCLES GRBM_GUI_ACTIVE "
          "SQ_INSTS_VALU_MFMA_MOPS_F64 SQ_INSTS_VALU_MFMA_MOPS_BF16 -- python3 bench.py --config c4 --steps 1 --quick)")
SIMD_NUM = 1024          # 256 CUs x 4 SIMDs


def main_c4(out=None):
    """profiles/gj_sweep_c4_trace.txt (per-kernel statistics of both plans over one warm-up and one
    timed inverse each) and profiles/gj_sweep_c4_mfma.txt (per kernel: MFMA busy cycles over
    GRBM_GUI_ACTIVE x 1024 SIMDs, and the MFMA flops the MOPS counters count)."""
    out = out or os.path.join(ROOT, "profile_out", "prof_c4")
    stats = list(csv.DictReader(open(os.path.join(out, "trace", "c4_kernel_stats.csv"))))
    bench = None
    with open(os.path.join(out, "trace", "..", "trace.log")) as f:
        for line in f:
            if line.startswith('{"metric"'):
                bench = json.loads(line)
    lines = ["# Kernel trace of the macro-atom inverse (2,000 cells N = I - Q, n in {256, 512, 1024})",
             f"command: {C4_CMD}", "",
             "Each plan (fp64: FP64 MFMA sweep; mixed: BF16x3 sweep + FP64 refinement) ran one warm-up and "
             "one timed inverse: per-batch kernel time = total / 2.", "",
             f"{'kernel':52s} {'calls':>6s} {'avg us':>10s} {'total ms':>10s} {'per batch ms':>12s} {'%':>6s}"]
    for r in stats[:18]:
        if r["Name"].startswith("__amd"):
            continue
        lines.append(f"{r['Name'][:52]:52s} {int(r['Calls']):6d} {float(r['AverageNs']) / 1e3:10.1f} "
                     f"{float(r['TotalDurationNs']) / 1e6:10.2f} {float(r['TotalDurationNs']) / 2e6:12.2f} "
                     f"{float(r['Percentage']):6.2f}")
    if bench:
        for mode, d in bench["modes"].items():
            lines.append(f"bench.py {mode}: {d['ms_per_step']} ms per batch (HIP events), {d['inverses_per_s']} "
                         f"inverses/s, {d['fp64_equivalent_tflops']} TFLOP/s of 2 n^3 per inverse")
    with open(os.path.join(ROOT, "profiles", "gj_sweep_c4_trace.txt"), "w") as f:
        f.write("\n".join(lines) + "\n")
    per, dur, seen = {}, {}, set()
    for r in csv.DictReader(open(os.path.join(out, "mfma", "c4_counter_collection.csv"))):
        k = r["Kernel_Name"][:52]
        d = per.setdefault(k, {})
        d[r["Counter_Name"]] = d.get(r["Counter_Name"], 0.0) + float(r["Counter_Value"])
        if r["Dispatch_Id"] not in seen:
            seen.add(r["Dispatch_Id"])
            dur[k] = dur.get(k, 0.0) + (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) * 1e-9
            d["dispatches"] = d.get("dispatches", 0) + 1
    ml = ["# MFMA counters of the macro-atom inverse kernels (a tenth of the C4 batch: 200 cells, both plans)",
          f"command: {C4_CMD}", "",
          "GRBM_GUI_ACTIVE per dispatch is the sum over the 8 XCDs (it reads 8 x clock x duration), so",
          "MFMA busy = SQ_VALU_MFMA_BUSY_CYCLES / (GRBM_GUI_ACTIVE / 8 x 1024 SIMDs); MFMA flop rates =",
          "MOPS x 512 over the dispatches' durations, against the dense peaks 78.6 TFLOP/s (FP64) and",
          "2500 TFLOP/s (BF16).", "",
          f"{'kernel':52s} {'disp':>5s} {'ms':>8s} {'MFMA busy %':>11s} {'F64 TF/s':>9s} {'% F64':>6s} "
          f"{'BF16 TF/s':>9s} {'% BF16':>7s}"]
    for k, d in sorted(per.items(), key=lambda kv: -kv[1].get("SQ_VALU_MFMA_BUSY_CYCLES", 0.0)):
        busy = d.get("SQ_VALU_MFMA_BUSY_CYCLES", 0.0)
        if busy == 0:
            continue
        act = d.get("GRBM_GUI_ACTIVE", 0.0) / 8.0
        t = dur[k]
        f64 = d.get("SQ_INSTS_VALU_MFMA_MOPS_F64", 0.0) * 512 / t / 1e12
        b16 = d.get("SQ_INSTS_VALU_MFMA_MOPS_BF16", 0.0) * 512 / t / 1e12
        ml.append(f"{k:52s} {int(d['dispatches']):5d} {t * 1e3:8.2f} {100 * busy / (act * SIMD_NUM):11.2f} "
                  f"{f64:9.2f} {100 * f64 / 78.6:6.2f} {b16:9.2f} {100 * b16 / 2500:7.2f}")
    with open(os.path.join(ROOT, "profiles", "gj_sweep_c4_mfma.txt"), "w") as f:
        f.write("\n".join(ml) + "\n")
    print("\n".join(lines + [""] + ml))


STALL_CMD = ("python_amd/profile_stalls.sh  (rocprofv3 --pmc <8 SQ counters>, twice, -- python3 bench.py "
             "--roofline-only --no-cpu)")


def main_stalls(out=None):
    """profiles/stall_counters_uneven10k.txt: per kernel of the bench workload, the fraction of
    wave cycles spent waiting (SQ_WAIT_ANY: on memory, LDS, barriers), waiting to issue
    (SQ_WAIT_INST_ANY), issuing VALU and LDS instructions, and LDS bank conflicts per LDS cycle."""
    out = out or os.path.join(ROOT, "profile_out", "stalls")
    agg = {}
    for sub in ("p1", "p2"):
        with open(os.path.join(out, sub, "st_counter_collection.csv")) as f:
            for r in csv.DictReader(f):
                k = r["Kernel_Name"]
                if not (k.startswith("void k_") or k.startswith("k_")):
                    continue
                k = k.replace("void ", "").split("(")[0]
                agg.setdefault(k, {})
                agg[k][r["Counter_Name"]] = agg[k].get(r["Counter_Name"], 0.0) + float(r["Counter_Value"])
    lines = ["# Stall counters of the bench workload's kernels (uneven 10k-cell batch, size classes in isolation)",
             f"command: {STALL_CMD}", "",
             "Fractions of SQ_WAVE_CYCLES (summed over the kernel's dispatches): wait = SQ_WAIT_ANY (a wave",
             "waiting on memory, LDS or a barrier), issue-wait = SQ_WAIT_INST_ANY (ready, not issued), valu =",
             "SQ_ACTIVE_INST_VALU, lds = SQ_ACTIVE_INST_LDS; bank = SQ_LDS_BANK_CONFLICT per LDS-active cycle.", "",
             f"{'kernel':44s} {'wait':>6s} {'issue-wait':>10s} {'valu':>6s} {'lds':>6s} {'bank':>6s}"]
    for k in sorted(agg, key=lambda k: -agg[k].get("SQ_WAVE_CYCLES", 0.0)):
        d = agg[k]
        wc = d.get("SQ_WAVE_CYCLES", 0.0)
        if wc <= 0:
            continue
        lds = d.get("SQ_ACTIVE_INST_LDS", 0.0)
        lines.append(f"{k[:44]:44s} {d.get('SQ_WAIT_ANY', 0.0) / wc:6.2f} {d.get('SQ_WAIT_INST_ANY', 0.0) / wc:10.2f} "
                     f"{d.get('SQ_ACTIVE_INST_VALU', 0.0) / wc:6.2f} {lds / wc:6.2f} "
                     f"{d.get('SQ_LDS_BANK_CONFLICT', 0.0) / lds if lds else 0.0:6.3f}")
    os.makedirs(os.path.join(ROOT, "profiles"), exist_ok=True)
    with open(os.path.join(ROOT, "profiles", "stall_counters_uneven10k.txt"), "w") as f:
        f.write("\n".join(lines) + "\n")
    print("\n".join(lines))

ION_CMD = ("python_amd/profile_ion.sh  (rocprofv3 --kernel-trace --stats --output-format csv -- python3 bench.py "
           "--config ion --steps 1 --warmup 1 --no-cpu)")


def main_ion(out=None):
    """profiles/ion_iteration_uneven10k_trace.txt: the last populations() call of the ion bench
    (every cell's whole n_e iteration): per-kernel dispatch counts and times, and the wall time of
    an iteration (assembly start to update end) in the first iterations, with most cells active,
    and in the tail, where a few cells are left and the order 175 < n <= 256 route decides."""
    out = out or os.path.join(ROOT, "profile_out", "ion")
    path = None
    for dp, _, fs in os.walk(os.path.join(out, "trace")):
        for f in fs:
            if f.endswith("kernel_trace.csv"):
                path = os.path.join(dp, f)
    rows = []
    with open(path) as f:
        for r in csv.DictReader(f):
            rows.append((int(r["Start_Timestamp"]), int(r["End_Timestamp"]), r["Kernel_Name"].replace("void ", "").split("(")[0]))
    rows.sort()
    last_init = max(i for i, r in enumerate(rows) if r[2] == "k_ion_init")
    call = [r for r in rows[last_init:] if r[2].startswith("k_")]
    t_call = (max(r[1] for r in call) - call[0][0]) / 1e6
    stats = {}
    for s0, e0, k in call:
        d = stats.setdefault(k, [0, 0.0])
        d[0] += 1
        d[1] += (e0 - s0) / 1e6
    # iterations: from each assembly to the update that follows it
    its = []
    cur = None
    for s0, e0, k in call:
        if k == "k_ion_assemble":
            cur = {"start": s0, "big": 0.0}
        elif cur is not None and k.startswith(("k_lookahead_f64", "k_blocked_f64<256", "k_bsub_f64<4>")):
            cur["big"] += (e0 - s0) / 1e6
        elif cur is not None and k == "k_ion_update":
            cur["end"] = e0
            its.append(cur)
            cur = None
    wall = [(i["end"] - i["start"]) / 1e6 for i in its]
    big = [i["big"] for i in its]
    lines = ["# Lock-step ionization iteration (10k uneven cells, n in {12..256}, seed 6): one populations() call",
             f"command: {ION_CMD}", "",
             f"call: {t_call:.2f} ms over {len(its)} iterations (assembly start to update end of each, graph",
             "replays of four iterations with a host read of the active count between them)", "",
             f"{'kernel':34s} {'calls':>6s} {'total ms':>9s} {'avg us':>8s}"]
    for k in sorted(stats, key=lambda k: -stats[k][1]):
        c, t = stats[k]
        lines.append(f"{k[:34]:34s} {c:6d} {t:9.2f} {1e3 * t / c:8.1f}")
    def band(lo, hi):
        w = wall[lo:hi]
        b = big[lo:hi]
        return (f"iterations {lo + 1}..{min(hi, len(wall))}: wall {sum(w) / len(w):.3f} ms per iteration, of which the "
                f"175 < n <= 256 route (factor kernels + k_bsub_f64<4>) {sum(b) / len(b):.3f} ms") if w else ""
    lines += ["", band(0, 6), band(6, 40), band(40, len(wall))]
    os.makedirs(os.path.join(ROOT, "profiles"), exist_ok=True)
    with open(os.path.join(ROOT, "profiles", "ion_iteration_uneven10k_trace.txt"), "w") as f:
        f.write("\n".join(lines) + "\n")
    print("\n".join(lines))


TL_CMD = ("python_amd/profile_timeline.sh  (rocprofv3 --kernel-trace --output-format csv -- python3 -m "
          "python_amd.profile_summary run_timeline)")


def run_timeline():
    """The traced workload: the bench's headline plan (uneven 10k-cell batch, mixed mode, autotuned,
    graph replays), three warm-up solves, a spin kernel as a marker, then two solves."""
    import torch
    import python_amd as hip
    from python_amd.generators import uneven
    b = uneven()
    A, B = hip.to_device(b)
    X = torch.zeros_like(B)
    codes = torch.empty(b.ncells, dtype=torch.int32, device=B.device)
    plan = hip.Plan.for_batch(b, mode="mixed", graph=True)
    plan.autotune(A, B, X, codes)
    for _ in range(3):
        plan.solve(A, B, X, codes)
    torch.cuda.synchronize()
    torch.cuda._sleep(1000000)
    torch.cuda.synchronize()
    for _ in range(2):
        plan.solve(A, B, X, codes)
    torch.cuda.synchronize()
    print("groups", [(g["nmax"], g["count"], g["fp64_kernel"]) for g in plan.describe()["groups"]], flush=True)


def main_timeline(out=None):
    """profiles/c2_step_timeline_uneven10k.txt: the kernels of the last traced solve of the headline
    plan, with their queue, start and end (us from the step's first kernel), grid, registers and
    static LDS, as the profiler saw them."""
    out = out or os.path.join(ROOT, "profile_out", "timeline")
    path = None
    for dp, _, fs in os.walk(os.path.join(out, "trace")):
        for f in fs:
            if f.endswith("kernel_trace.csv"):
                path = os.path.join(dp, f)
    with open(path) as f:
        rows = sorted(csv.DictReader(f), key=lambda r: int(r["Start_Timestamp"]))
    mark = max(i for i, r in enumerate(rows) if "spin" in r["Kernel_Name"] or "sleep" in r["Kernel_Name"])
    after = rows[mark + 1:]
    # the second marked solve starts with the second launch of the step's first kernel
    first = after[0]["Kernel_Name"]
    starts = [i for i, r in enumerate(after) if r["Kernel_Name"] == first]
    step = after[starts[1]:] if len(starts) > 1 else after
    t0 = int(step[0]["Start_Timestamp"])
    t_end = max(int(r["End_Timestamp"]) for r in step)
    lines = ["# One solve of the headline plan (uneven 10k-cell batch, mixed mode, autotuned: every class in FP64)",
             f"command: {TL_CMD}", "",
             f"step: {(t_end - t0) / 1e3:.1f} us from the first kernel's start to the last kernel's end "
             "(profiled; queue = HIP stream)", "",
             "VGPR as the trace reports it (units of two registers); LDS the static part only (the factor",
             "kernels' dynamic LDS: k_blocked_f64<256> 72 KB per workgroup at n = 256, k_wave_f64<64> 9.4 KB,",
             "k_wave_f64<32, 2> 9.6 KB, k_wave_f64<16, 4> 11 KB).", "",
             f"{'start':>8s} {'end':>8s} {'dur':>7s} {'queue':>5s} {'grid':>6s} {'vgpr':>5s} {'lds':>6s}  kernel"]
    for r in step:
        s0 = (int(r["Start_Timestamp"]) - t0) / 1e3
        e0 = (int(r["End_Timestamp"]) - t0) / 1e3
        grid = int(r["Grid_Size_X"]) // max(1, int(r["Workgroup_Size_X"]))
        name = r["Kernel_Name"].replace("void ", "").split("(")[0]
        lines.append(f"{s0:8.1f} {e0:8.1f} {e0 - s0:7.1f} {r['Queue_Id']:>5s} {grid:6d} {r['VGPR_Count']:>5s} "
                     f"{r['LDS_Block_Size']:>6s}  {name}")
    os.makedirs(os.path.join(ROOT, "profiles"), exist_ok=True)
    with open(os.path.join(ROOT, "profiles", "c2_step_timeline_uneven10k.txt"), "w") as f:
        f.write("\n".join(lines) + "\n")
    print("\n".join(lines))


if __name__ == "__main__":
    if len(sys.argv) > 1 and sys.argv[1] == "c4":
        main_c4(sys.argv[2] if len(sys.argv) > 2 else None)
    elif len(sys.argv) > 1 and sys.argv[1] == "stalls":
        main_stalls(sys.argv[2] if len(sys.argv) > 2 else None)
    elif len(sys.argv) > 1 and sys.argv[1] == "ion":
        main_ion(sys.argv[2] if len(sys.argv) > 2 else None)
    elif len(sys.argv) > 1 and sys.argv[1] == "run_timeline":
        run_timeline()
    elif len(sys.argv) > 1 and sys.argv[1] == "timeline":
        main_timeline(sys.argv[2] if len(sys.argv) > 2 else None)
    else:
        main(sys.argv[1] if len(sys.argv) > 1 else None)
