#!/usr/bin/env python3
"""Throughput of the per-cell dense solve on one MI355X, on the 10,000-cell uneven batch.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-plain] [--roofline-only]

One "step" = one solve of the whole batch: 10,000 ionization-balance cells, n in
{12, 30, 64, 96, 128, 200, 256} with counts 3000/2500/1500/1000/1000/500/500, seed 2, inputs
resident in HBM.  ``value`` is the batched mixed-precision solver (equilibrated FP32 LU + FP64
refinement + FP64 re-solve of handed-back cells, size classes routed by the autotuner, one graph
launch per step) in solves per second; the line also carries the batched FP64 GSL-order solver,
the plain per-matrix FP64 launches (one launch per cell, as the reference's per-call GPU path),
the per-cell error against the FP64 oracle, the roofline of the dominant kernel and the CPU
baseline (the oracle, one worker per host core).

Roofline: the dominant kernel is the kernel instantiation with the largest total time over the
plan's size classes.  Its launches are timed in isolation (HIP events on the plan's stream,
``hipla_plan_time_groups``), ``achieved`` = the algorithmic bytes / flops of the cells those
launches solve divided by their total time, and ``kernel_avg_ms`` the mean launch duration.
``--roofline-only`` runs just the plan set-up and this measurement (plus one elementwise kernel
that reads and writes 512 MiB, to calibrate HBM byte counters) and prints the roofline object,
for profiler runs: the dominant kernel is then launched only by the timed launches.  ``traffic`` is
read from profiles/traffic.json when it names the same kernel (rocprofv3 FETCH_SIZE +
WRITE_SIZE per launch, scaled by the copy calibration), else null.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
FP32_PEAK_TFLOPS = 157.3
BF16_PEAK_TFLOPS = 2500.0


def _cores():
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def measure_roofline(plan, A, B, reps=10):
    """Dominant kernel of an FP64-order plan, from isolated launches of its size classes.

    Every size class of ``plan`` is launched ``reps`` times on its own (HIP events on the plan's
    stream, nothing else launched); the kernel instantiation with the largest total time is the
    dominant one.  Returns (roofline dict, per-class list)."""
    import torch
    X = torch.empty_like(B)
    codes = torch.empty(plan.ncells, dtype=torch.int32, device=B.device)
    desc = plan.describe()
    gms = plan.time_groups(A, B, X, codes, reps=reps, classes_only=True)
    by_kernel = {}
    classes = []
    for g, t in zip(desc["groups"], gms):
        name = g["fp64_kernel"]
        k = by_kernel.setdefault(name, {"ms": 0.0, "bytes": 0.0, "flops": 0.0, "launches": 0, "classes": []})
        k["ms"] += float(t)
        k["bytes"] += float(g["bytes"])
        k["flops"] += float(g["flops"])
        k["launches"] += 1
        k["classes"].append(f"n<={g['nmax']} x{g['count']}")
        classes.append({"kernel": name, "nmax": g["nmax"], "count": g["count"], "ms": round(float(t), 4),
                        "bytes": float(g["bytes"]), "flops": float(g["flops"])})
    dom = max(by_kernel, key=lambda nm: by_kernel[nm]["ms"])
    dk = by_kernel[dom]
    bw = dk["bytes"] / (dk["ms"] * 1e-3) / 1e9
    fl = dk["flops"] / (dk["ms"] * 1e-3) / 1e12
    # the FP64 kernels run on the vector ALUs (GSL's order: separately rounded products, no
    # matrix cores), so the compute roof is the FP64 vector peak
    if bw / HBM_PEAK_GBS >= fl / FP64_PEAK_TFLOPS:
        roof = {"bound": "hbm", "achieved": round(bw, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(bw / HBM_PEAK_GBS, 5)}
    else:
        roof = {"bound": "fp64", "achieved": round(fl, 4), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": round(fl / FP64_PEAK_TFLOPS, 5)}
    roof["traffic"] = None
    roof["kernel"] = dom
    roof["kernel_avg_ms"] = round(dk["ms"] / dk["launches"], 4)
    roof["launches_per_step"] = dk["launches"]
    roof["size_classes"] = dk["classes"]
    roof["algorithmic_bytes_per_launch"] = round(dk["bytes"] / dk["launches"])
    roof["algorithmic_flops_per_launch"] = round(dk["flops"] / dk["launches"])
    roof["hbm_achieved_gbs"] = round(bw, 2)
    roof["fp64_achieved_tflops"] = round(fl, 4)
    tpath = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            tr = json.load(f)
        if tr.get("kernel") == dom:
            roof["traffic"] = tr.get("hbm_bytes_per_launch")
    return roof, classes


def _timed_ms(fn, steps, warmup):
    """Average ms per call of fn() between HIP events on torch's current stream."""
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(steps):
        fn()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / steps


def _tune(plan, A, B):
    """hipla_plan_autotune on the bench inputs (kernel shapes per size class)."""
    import torch
    X = torch.empty_like(B)
    codes = torch.empty(plan.ncells, dtype=torch.int32, device=B.device)
    plan.autotune(A, B, X, codes)
    torch.cuda.synchronize()


def _emit(out):
    print(json.dumps(out), flush=True)
    return 0


def bench_c1(args):
    """The plain small case: 900 cells of the n = 30 ionization matrix (H, He, C, N, O + slack,
    block-tridiagonal), seed 1.  The batched FP64 GSL-order solve (one kernel launch per step)
    against one launch per matrix and the CPU oracle on the same host, values and codes checked
    against the oracle in every cell."""
    import torch
    import oracle
    import python_amd as pa
    from python_amd.generators import plain_small
    batch = plain_small()
    A, B = pa.to_device(batch)
    plan = pa.Plan.for_batch(batch, mode="fp64", graph=True)
    X = torch.zeros_like(B)
    codes = torch.empty(batch.ncells, dtype=torch.int32, device=B.device)
    ms = _timed_ms(lambda: plan.solve(A, B, X, codes), max(args.steps, 50), args.warmup)
    per = pa.Plan.for_batch(batch, mode="fp64", graph=False, per_cell=True)
    Xp = torch.zeros_like(B)
    ms_per = _timed_ms(lambda: per.solve(A, B, Xp, codes), 3, 1)
    oracle.build()
    cores = _cores()
    Xo, co = oracle.solve_batch(batch.n, batch.offA, batch.offV, batch.A, batch.B, nthreads=cores)
    plan.solve(A, B, X, codes)
    torch.cuda.synchronize()
    equal = bool(np.array_equal(X.cpu().numpy(), Xo) and np.array_equal(codes.cpu().numpy(), co))
    reps_c, total = 0, 0.0
    while total < 10.0 and reps_c < 100000:
        t0 = time.perf_counter()
        oracle.solve_batch(batch.n, batch.offA, batch.offV, batch.A, batch.B, nthreads=cores)
        total += time.perf_counter() - t0
        reps_c += 1
    reps_1, total_1 = 0, 0.0
    while total_1 < 5.0 and reps_1 < 2000:
        t0 = time.perf_counter()
        oracle.solve_batch(batch.n, batch.offA, batch.offV, batch.A, batch.B, nthreads=1)
        total_1 += time.perf_counter() - t0
        reps_1 += 1
    byts, flops = float(batch.bytes_solve), batch.flops_solve
    bw = byts / (ms * 1e-3) / 1e9
    return _emit({
        "metric": "solves/s, 900 ionization matrices n = 30, FP64 GSL order", "unit": "solves/s",
        "value": round(batch.ncells / (ms * 1e-3), 1), "ms_per_step": round(ms, 4), "n_gpus": 1, "dtype": "f64",
        "data": "synthetic (seeded block-tridiagonal ionization matrices, seed 1)",
        "config": {"workload": "plain_small: 900 cells, n = 30 (H2 He3 C7 N8 O9 + slack)"},
        "equal_to_oracle_values_and_codes": equal,
        "plain_fp64_per_matrix": {"solves_per_s": round(batch.ncells / (ms_per * 1e-3), 1),
                                  "ms_per_step": round(ms_per, 3)},
        "roofline": {"bound": "hbm", "achieved": round(bw, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(bw / HBM_PEAK_GBS, 5), "traffic": None,
                     "kernel": plan.describe()["groups"][0]["fp64_kernel"] + " (one launch per step)",
                     "algorithmic_bytes_per_launch": round(byts), "algorithmic_flops_per_launch": round(flops)},
        "cpu_baseline": {"value": round(reps_c * batch.ncells / total, 1), "unit": "solves/s", "cores": cores,
                         "kind": "port", "sample": f"the 900-cell batch x {reps_c} ({total:.1f} s, one OpenMP "
                                                   "worker per core)"},
        "cpu_single_core": {"value": round(reps_1 * batch.ncells / total_1, 1), "unit": "solves/s", "cores": 1,
                            "kind": "port", "sample": f"the 900-cell batch x {reps_1} ({total_1:.1f} s)"}})


def bench_c3(args):
    """Wide-range rates (inside 1e-20..1e4) with 2-norm condition numbers 1e6..1e14 (histogram
    logged), the same 10,000-cell size mix, seed 3: the mixed solver with every size class on the
    FP32 path (no tuning), then autotuned; per-cell error against the oracle, refinement-step
    histogram, fallback rate."""
    import torch
    import oracle
    import python_amd as pa
    from python_amd.generators import uneven
    batch = uneven(seed=3, wide=True, scale=0.1 if args.quick else 1.0)
    A, B = pa.to_device(batch)
    oracle.build()
    Xo, co = oracle.solve_batch(batch.n, batch.offA, batch.offV, batch.A, batch.B, nthreads=_cores())
    out = {"metric": "solves/s, wide-range rate matrices (1e-20..1e4, kappa 1e6..1e14), mixed precision",
           "unit": "solves/s", "n_gpus": 1,
           "data": "synthetic (seeded ionization-balance generator, wide rates, kappa targeted, seed 3)",
           "config": {"workload": f"uneven_wide_10k: {batch.ncells} cells, n in {{12..256}}"}, "modes": {}}
    for mode, tune in (("fp32_refined_untuned", False), ("autotuned", True)):
        plan = pa.Plan.for_batch(batch, mode="mixed", graph=True)
        X = torch.zeros_like(B)
        codes = torch.empty(batch.ncells, dtype=torch.int32, device=B.device)
        st = torch.zeros(batch.ncells * 24, dtype=torch.uint8, device=B.device)   # hipla_cell_stats records
        if tune:
            plan.autotune(A, B, X, codes)
        ms = _timed_ms(lambda: plan.solve(A, B, X, codes, st), args.steps, args.warmup)
        Xh, ch = X.cpu().numpy(), codes.cpu().numpy()
        stats = pa.stats_to_numpy(st)
        rel = np.zeros(batch.ncells)
        for c in np.where(co != 4)[0]:
            o, n = batch.offV[c], batch.n[c]
            rel[c] = np.abs(Xh[o:o + n] - Xo[o:o + n]).max() / max(np.abs(Xo[o:o + n]).max(), 1e-300)
        fb = stats["fallback"]
        out["modes"][mode] = {
            "solves_per_s": round(batch.ncells / (ms * 1e-3), 1), "ms_per_step": round(ms, 4),
            "max_rel_err_vs_fp64_oracle": float(rel.max()), "cells_within_1e-8": int((rel <= 1e-8).sum()),
            "codes_equal_oracle": bool(np.array_equal(ch, co)),
            "fallback_rate": round(float((fb == 1).sum()) / batch.ncells, 4),
            "fp64_direct_cells": int((fb == 2).sum()),
            "refine_steps_hist": np.bincount(stats["refine_steps"][fb == 0], minlength=1).tolist(),
            "per_size": {int(n): {"cells": int((batch.n == n).sum()),
                                  "fallback": int(((fb == 1) & (batch.n == n)).sum()),
                                  "max_rel_err": float(rel[batch.n == n].max())} for n in np.unique(batch.n)}}
        print(f"[bench c3] {mode}: {ms:.3f} ms", file=sys.stderr, flush=True)
        plan.close()
    out["value"] = out["modes"]["autotuned"]["solves_per_s"]
    out["singular_cells_code4"] = int((co == 4).sum())
    out["kappa_hist"] = batch.meta["kappa_hist"]
    out["log10_kappa_range"] = [round(v, 3) for v in batch.meta["log10_kappa_range"]]
    out["abs_entry_range"] = batch.meta["abs_entry_range"]
    return _emit(out)


def bench_c4(args):
    """Macro-atom inversion: 2,000 cells N = I - Q, n in {256, 512, 1024} (thirds), seed 4.  Two
    plans on the same cells: FP64 (blocked Gauss-Jordan sweep with FP64 MFMA block products) and
    mixed (the sweep in FP32 with split BF16x3 MFMA products, then two FP64 refinement steps);
    ``value`` is the faster one.  Relative Frobenius error against the oracle on two cells per
    order.  Roofline: algorithmic flops 2 n^3 per inverse over the step time, against the FP64
    peak (both plans' dominant work is FP64: the sweep's block products, or the residual GEMMs)."""
    import torch
    import oracle
    import python_amd as pa
    from python_amd.generators import macro_inverse
    t0 = time.time()
    batch = macro_inverse(ncells=200 if args.quick else 2000)
    print(f"[bench c4] generated {batch.ncells} cells in {time.time() - t0:.1f}s", file=sys.stderr, flush=True)
    A, _ = pa.to_device(batch)
    flops = float(sum(2.0 * n ** 3 for n in batch.n))
    byts = float(sum(16.0 * n * n for n in batch.n))
    oracle.build()
    refs = {}
    for nsz in np.unique(batch.n):
        for c in np.where(batch.n == nsz)[0][:2]:
            a, _ = batch.cell(c)
            refs[int(c)] = oracle.invert_matrix(a)[1]
    modes = {}
    Inv = torch.empty_like(A)
    for mode in ("fp64", "mixed"):
        plan = pa.Plan.for_batch(batch, mode=mode, graph=True)
        codes = torch.empty(batch.ncells, dtype=torch.int32, device=A.device)
        steps = max(1, min(args.steps, 5))
        ms = _timed_ms(lambda: plan.invert(A, Inv, codes), steps, 1)
        Ih, ch = Inv.cpu().numpy(), codes.cpu().numpy()
        errs = {}
        for c, ref in refs.items():
            nsz = int(batch.n[c])
            got = Ih[batch.offA[c]:batch.offA[c] + nsz * nsz].reshape(nsz, nsz)
            errs.setdefault(nsz, []).append(float(np.linalg.norm(got - ref) / np.linalg.norm(ref)))
        tf = flops / (ms * 1e-3) / 1e12
        # MFMA work the plan issues (64 x 64 x 64 block products: nb (nb^2 - 1) per sweep of an order
        # n = 64 nb; mixed adds two refinement steps of nb^2 blocks with K = n: one FP64 residual,
        # one BF16x3 correction); BF16x3 issues three BF16 products per block product
        blk = 2.0 * 64 ** 3
        sweep = float(sum((n // 64) * ((n // 64) ** 2 - 1) * blk for n in batch.n))
        refine = float(sum(2 * (n // 64) ** 2 * 2.0 * 64 * 64 * n for n in batch.n))
        f64_mfma = sweep if mode == "fp64" else refine
        bf16_mfma = 0.0 if mode == "fp64" else 3.0 * (sweep + refine)
        mfma = {"fp64_mfma_tflops": round(f64_mfma / (ms * 1e-3) / 1e12, 3),
                "fp64_mfma_frac": round(f64_mfma / (ms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS, 5),
                "bf16_mfma_tflops": round(bf16_mfma / (ms * 1e-3) / 1e12, 3),
                "bf16_mfma_frac": round(bf16_mfma / (ms * 1e-3) / 1e12 / BF16_PEAK_TFLOPS, 5)}
        modes[mode] = {"inverses_per_s": round(batch.ncells / (ms * 1e-3), 2), "ms_per_step": round(ms, 3),
                       "mfma_whole_step": mfma,
                       "steps": steps, "codes_nonzero": int((ch != 0).sum()), "rel_frobenius_err_sample": errs,
                       "max_rel_frobenius_err_sample": max(max(v) for v in errs.values()),
                       "fp64_equivalent_tflops": round(tf, 3)}
        print(f"[bench c4] {mode}: {ms:.2f} ms", file=sys.stderr, flush=True)
        plan.close()
    best = min(modes, key=lambda m: modes[m]["ms_per_step"])
    ms = modes[best]["ms_per_step"]
    tf = flops / (ms * 1e-3) / 1e12
    return _emit({
        "metric": "inverses/s, macro-atom matrices N = I - Q, n in {256,512,1024}", "unit": "inverses/s",
        "value": modes[best]["inverses_per_s"], "ms_per_step": ms, "n_gpus": 1,
        "dtype": "f64" if best == "fp64" else "bf16x3+f64", "data": "synthetic (seeded substochastic Q, seed 4)",
        "config": {"workload": f"macro_inverse: {batch.ncells} cells, n in {{256,512,1024}} thirds", "plan": best},
        "modes": modes,
        "roofline": {"bound": "fp64", "achieved": round(tf, 4), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(tf / FP64_PEAK_TFLOPS, 5), "traffic": None,
                     "kernel": "whole inverse (Gauss-Jordan sweep on the matrix cores)",
                     "hbm_achieved_gbs": round(byts / (ms * 1e-3) / 1e9, 2)}})


def bench_c5(args):
    """T_e gradient: 10,000 cells of n = 64 whose rates depend on T_e, seed 5.  A step is the GPU
    assembly of A(T_e) (hipla_te_assemble), the batched FP64 solve (factors kept), the adjoint
    A^T lambda = dL/dx and dL/dT_e straight from lambda (hipla_te_grad), behind
    torch.autograd; the same step through torch's assembly and chain rule over the dense dL/dA is
    timed beside it.  The gradient is checked against the FP64 oracle on a sample."""
    import torch
    import oracle
    from python_amd.adjoint import IonTe, loss_and_grad
    from python_amd.generators import ion_matrices_te, te_assemble
    rng = np.random.default_rng(5)
    ncells = 1000 if args.quick else 10000
    d = ion_matrices_te(rng, ncells, 64)
    w = rng.random((ncells, 64))
    model = IonTe(d)
    plan = model.plan()
    T = torch.from_numpy(d["te"]).cuda()
    wt = torch.from_numpy(w).cuda()
    res = {}

    def step_torch():
        res["out"] = loss_and_grad(model, plan, T, wt, fused=False)

    ms_torch = _timed_ms(step_torch, args.steps, args.warmup)

    def step():
        res["out"] = loss_and_grad(model, plan, T, wt)

    ms = _timed_ms(step, args.steps, args.warmup)
    L, g, codes = res["out"]
    g = g.cpu().numpy()
    oracle.build()
    A, dA, b = te_assemble(64, d["blocks"], d["params"], d["te"])
    worst = 0.0
    for c in range(0, ncells, max(1, ncells // 200)):
        LU, perm, _, _ = oracle.lu_decomp(A[c])
        _, x = oracle.lu_solve(LU, perm, b[c])
        _, lam = oracle.lu_solve_transpose(LU, perm, w[c])
        gref = -lam @ (dA[c] @ x)
        mag = np.abs(lam) @ (np.abs(dA[c]) @ np.abs(x))
        worst = max(worst, abs(gref - g[c]) / (1e-8 * abs(gref) + 1e-10 * mag))
    return _emit({
        "metric": "gradients/s, dL/dT_e through the adjoint, n = 64", "unit": "cells/s",
        "value": round(ncells / (ms * 1e-3), 1), "ms_per_step": round(ms, 4), "n_gpus": 1, "dtype": "f64",
        "data": "synthetic (T_e-dependent ionization matrices, seed 5)",
        "config": {"workload": f"te_gradient: {ncells} cells, n = 64, loss = sum_i w_i x_i"},
        "torch_assembly_and_chain_rule": {"cells_per_s": round(ncells / (ms_torch * 1e-3), 1),
                                          "ms_per_step": round(ms_torch, 4)},
        "gradient_error_over_bound_max": round(worst, 6),
        "gradient_bound": "1e-8 |g| + 1e-10 sum|lambda||dA||x| against the oracle on 200 sampled cells"})


def bench_ion(args):
    """The ionization balance of 10,000 cells in lock step (the uneven order mix 12..256, seed 6):
    one step = the whole n_e fixed point of every cell (hipla_ion_populations: GPU assembly of the
    rate matrices, FP64 GSL-order solves of the cells still iterating, floors / damping /
    convergence on the GPU) from the same starting densities.  Checked bit for bit against the
    oracle's per-cell restatement of matrix_ion_populations, which also gives the CPU baseline
    (one worker per host core)."""
    import torch
    import oracle
    from python_amd.generators import ion_iteration_batch
    from python_amd.ion import IonIteration
    t0 = time.time()
    batch = ion_iteration_batch(scale=0.1 if args.quick else 1.0)
    print(f"[bench ion] generated {batch.ncells} cells in {time.time() - t0:.1f}s", file=sys.stderr, flush=True)
    it = IonIteration(batch)
    steps = max(1, min(args.steps, 10))
    ms = _timed_ms(lambda: it.populations(), steps, max(1, args.warmup))
    dens, ne, st, iters, codes = (t.cpu().numpy() for t in it.populations())
    oracle.build()
    cores = _cores()
    t0 = time.perf_counter()
    wd, wne, wst, wit, wcodes = oracle.ion_populations_batch(batch, nthreads=cores)
    t_cpu = time.perf_counter() - t0
    reps = 1
    while t_cpu < 10.0 and reps < 50 and not args.no_cpu:
        t1 = time.perf_counter()
        oracle.ion_populations_batch(batch, nthreads=cores)
        t_cpu += time.perf_counter() - t1
        reps += 1
    ok = wst == 0
    equal = bool(np.array_equal(st, wst) and np.array_equal(iters, wit) and np.array_equal(codes, wcodes)
                 and np.array_equal(ne[ok], wne[ok]) and np.array_equal(dens.view(np.uint64), wd.view(np.uint64)))
    # solves per cell: iterations + 1 for a converged cell, the iteration count for MAXITERATIONS,
    # iterations + 1 for a singular one
    nsolve = np.where((wst != 0) & (wit == 200), 200, wit + 1)
    n = batch.n.astype(np.float64)
    solve_bytes = float((nsolve * 8.0 * (n * n + 2 * n)).sum())
    # assembly: the n x n matrix and b written, the nine per-ion arrays read, per iteration
    asm_bytes = float((nsolve * 8.0 * (n * n + n + 9 * n)).sum())
    byts = solve_bytes + asm_bytes
    bw = byts / (ms * 1e-3) / 1e9
    return _emit({
        "metric": "ionization balances/s, lock-step n_e iteration (matrix_ion_populations) on 10k uneven cells",
        "unit": "cells/s", "value": round(batch.ncells / (ms * 1e-3), 1), "ms_per_step": round(ms, 3), "steps": steps,
        "n_gpus": 1, "dtype": "f64", "data": "synthetic (seeded per-ion rate coefficients, seed 6)",
        "config": {"workload": f"ion_iteration: {batch.ncells} cells, n in {{12..256}}, iterations to convergence"},
        "equal_to_oracle": equal, "solves_per_s": round(float(nsolve.sum()) / (ms * 1e-3), 1),
        "solves_per_step": int(nsolve.sum()), "iterations_max": int(wit.max()),
        "cells_failed": int((wst != 0).sum()), "iterations_hist": np.bincount(wit, minlength=1).tolist()[:30],
        "roofline": {"bound": "hbm", "achieved": round(bw, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(bw / HBM_PEAK_GBS, 5), "traffic": None,
                     "kernel": "whole iteration (assembly + FP64 solves + update), one step",
                     "algorithmic_bytes_per_step": round(byts)},
        "cpu_baseline": {"value": round(reps * batch.ncells / t_cpu, 1), "unit": "cells/s", "cores": cores,
                         "kind": "port", "sample": f"the whole batch x {reps} ({t_cpu:.1f} s, one OpenMP worker "
                                                   "per core)"}})


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--quick", action="store_true", help="1/10 of the batch (development only)")
    ap.add_argument("--roofline-only", action="store_true",
                    help="plan set-up, graph timing and the dominant-kernel measurement only (profiler runs)")
    ap.add_argument("--config", default="c2", choices=["c1", "c2", "c3", "c4", "c5", "ion"],
                    help="c1: 900 x n = 30; c2: the headline batch (default); c3: wide-range rates, "
                         "refinement statistics; "
                         "c4: macro-atom inversion; c5: T_e gradient (forward solve + adjoint); "
                         "ion: lock-step n_e iteration of the ionization balance")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build runs on a single MI355X (the cell batch would shard "
              "across GPUs by cumulative flops with one all-gather of the solutions; not run here)",
              file=sys.stderr)
        return 2

    import torch
    import python_amd as pa
    from python_amd.generators import uneven

    if args.config != "c2":
        return {"c1": bench_c1, "c3": bench_c3, "c4": bench_c4, "c5": bench_c5, "ion": bench_ion}[args.config](args)

    t_gen = time.time()
    batch = uneven(scale=0.1 if args.quick else 1.0)
    print(f"[bench] generated {batch.ncells} cells in {time.time() - t_gen:.1f}s", file=sys.stderr, flush=True)
    dev = torch.device("cuda:0")
    A, B = pa.to_device(batch, dev)
    torch.cuda.synchronize()

    if args.roofline_only:
        # only the isolated size-class launches of the FP64 plan, then one elementwise kernel that
        # reads and writes 512 MiB (the calibration for HBM byte counters); prints the roofline
        # object alone
        fp64 = pa.Plan.for_batch(batch, mode="fp64", graph=False)
        _tune(fp64, A, B)
        roof, classes = measure_roofline(fp64, A, B)
        src = torch.empty(64 * 1024 * 1024, dtype=torch.float64, device=dev).fill_(1.0)
        dst = torch.empty_like(src)
        torch.mul(src, 1.0, out=dst)      # one elementwise kernel: reads and writes 512 MiB
        torch.cuda.synchronize()
        print(json.dumps({"roofline": roof, "classes": classes, "calibration_copy_bytes": src.numel() * 8}),
              flush=True)
        return 0

    def timed(plan, steps, warmup, stats=False):
        X = torch.empty_like(B)
        codes = torch.empty(batch.ncells, dtype=torch.int32, device=dev)
        st = torch.zeros(batch.ncells * 24, dtype=torch.uint8, device=dev) if stats else None
        for _ in range(warmup):
            plan.solve(A, B, X, codes, st)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t0 = time.perf_counter()
        e0.record()
        for _ in range(steps):
            plan.solve(A, B, X, codes, st)
        e1.record()
        torch.cuda.synchronize()
        wall = (time.perf_counter() - t0) * 1e3 / steps
        return e0.elapsed_time(e1) / steps, wall, X, codes, st

    # ---- batched mixed (autotuned size-class routing), graph replay ----
    mixed = pa.Plan.for_batch(batch, mode="mixed", graph=True)
    Xt = torch.empty_like(B)
    ct = torch.empty(batch.ncells, dtype=torch.int32, device=dev)
    mixed.autotune(A, B, Xt, ct)
    torch.cuda.synchronize()
    ms_mixed, wall_mixed, Xm, cm, sm = timed(mixed, args.steps, args.warmup, stats=True)
    print(f"[bench] mixed: {ms_mixed:.3f} ms/step", file=sys.stderr, flush=True)

    # ---- batched FP64 GSL order (kernel shapes tuned per size class), graph replay ----
    fp64 = pa.Plan.for_batch(batch, mode="fp64", graph=True)
    _tune(fp64, A, B)
    ms_fp64, _, Xf, cf, _ = timed(fp64, args.steps, args.warmup)
    print(f"[bench] fp64 batched: {ms_fp64:.3f} ms/step", file=sys.stderr, flush=True)

    # ---- plain per-matrix FP64 (one launch per cell) ----
    ms_plain = None
    if not args.no_plain:
        plain = pa.Plan.for_batch(batch, mode="fp64", graph=False, per_cell=True)
        ms_plain, _, Xp, cp, _ = timed(plain, max(1, min(3, args.steps)), 1)
        print(f"[bench] fp64 per-matrix: {ms_plain:.3f} ms/step", file=sys.stderr, flush=True)
        plain.close()

    # ---- accuracy against the FP64 oracle (CPU, all host cores) ----
    import oracle
    oracle.build()
    cores = _cores()
    Xo, co = oracle.solve_batch(batch.n, batch.offA, batch.offV, batch.A, batch.B, nthreads=cores)
    Xm_h = Xm.cpu().numpy()
    Xf_h = Xf.cpu().numpy() if ms_fp64 is not None else None
    stats = pa.stats_to_numpy(sm)
    rel = np.zeros(batch.ncells)
    rel64 = np.zeros(batch.ncells)
    cm_h = cm.cpu().numpy()
    for c in range(batch.ncells):
        if co[c] == 4:          # singular in FP64: x is not written (solve_matrix code 4)
            continue
        o, n = batch.offV[c], batch.n[c]
        xo = Xo[o:o + n]
        sc = max(np.abs(xo).max(), 1e-300)
        rel[c] = np.abs(Xm_h[o:o + n] - xo).max() / sc
        if Xf_h is not None:
            rel64[c] = np.abs(Xf_h[o:o + n] - xo).max() / sc
    fb = stats["fallback"]
    steps_hist = np.bincount(stats["refine_steps"][fb == 0], minlength=1).tolist()

    # ---- per-class times of the mixed plan (isolated launches) ----
    desc = mixed.describe()
    groups = desc["groups"]
    gms = mixed.time_groups(A, B, Xt, ct, reps=5)

    # ---- roofline of the dominant kernel (isolated launches of the FP64 plan's size classes) ----
    roof, classes = measure_roofline(fp64, A, B)

    # ---- CPU baseline: the oracle on a bounded sample (the whole batch, repeated to >= 10 s) ----
    cpu = None
    if not args.no_cpu:
        reps_c, total = 0, 0.0
        while total < 10.0 and reps_c < 200:
            t0 = time.perf_counter()
            oracle.solve_batch(batch.n, batch.offA, batch.offV, batch.A, batch.B, nthreads=cores)
            total += time.perf_counter() - t0
            reps_c += 1
        cpu = {"value": round(reps_c * batch.ncells / total, 1), "unit": "solves/s", "cores": cores,
               "kind": "port", "sample": f"the full {batch.ncells}-cell batch x {reps_c} (GSL-order C restatement, "
                                         f"one OpenMP worker per core, {total:.1f} s)"}

    value = batch.ncells / (ms_mixed * 1e-3)
    out = {
        "metric": "solves/s, 10k uneven rate matrices ≤1e-8 rel err: batched-mixed vs plain FP64",
        "value": round(value, 1),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_mixed, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if int((fb == 2).sum()) == batch.ncells else "f32 LU + f64 refinement / f64",
        "data": "synthetic (seeded ionization-balance generator, seed 2)",
        "config": {"workload": "uneven_10k: 10000 cells, n in {12,30,64,96,128,200,256}, counts "
                               "3000/2500/1500/1000/1000/500/500, 1 RHS each, FP64 in/out",
                   "global_batch": batch.ncells, "parallelism": "single GPU, one graph launch per step"},
        "max_rel_err_vs_fp64_oracle": float(rel.max()),
        "cells_within_1e-8": int((rel <= 1e-8).sum()),
        "singular_cells_code4": int((co == 4).sum()),
        "codes_equal_oracle": bool(np.array_equal(cm_h, co)),
        "fallback_rate": round(float((fb == 1).sum()) / batch.ncells, 4),
        "fp64_direct_cells": int((fb == 2).sum()),
        "value_runs": ("the batched FP64 GSL-order kernels: the autotuner timed FP32 LU + FP64 refinement "
                       "against FP64 per size class and routed every class to FP64 (FP32 rounding of these "
                       "rate matrices drops their small rates; most FP32 factorisations are handed back)")
        if int((fb == 2).sum()) == batch.ncells else "mixed FP32 LU + FP64 refinement on the tuned classes",
        "refine_steps_hist": steps_hist,
        "batched_fp64": None if ms_fp64 is None else
        {"solves_per_s": round(batch.ncells / (ms_fp64 * 1e-3), 1), "ms_per_step": round(ms_fp64, 4),
         "max_rel_err_vs_oracle": float(rel64.max()), "codes_equal_oracle": bool(np.array_equal(cf.cpu().numpy(), co))},
        "plain_fp64_per_matrix": None if ms_plain is None else
        {"solves_per_s": round(batch.ncells / (ms_plain * 1e-3), 1), "ms_per_step": round(ms_plain, 3)},
        "speedup_vs_plain_fp64": None if ms_plain is None else round(ms_plain / ms_mixed, 2),
        "roofline": roof,
        "cpu_baseline": cpu,
        "groups": [{k: g2[k] for k in ("kernel", "fp64_kernel", "nmax", "count", "direct_fp64", "variant")} |
                   {"ms": round(float(t), 4)} for g2, t in zip(groups, gms)],
        "fp64_classes": classes,
        "fallback_pass_ms": round(float(gms[len(groups)]), 4),
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
