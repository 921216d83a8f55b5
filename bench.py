"""Throughput of the MI355X Plan7 filter path (MSV / Viterbi / Forward), batched vs plain.

A step is one pass of the hot path over the synthetic workload's pairs.  By default the
workload is the Pfam-scale set (20,000 profiles x 50,000 proteins, 6.1e13 MSV cells); the
timed stage is selected with --stage.  GCUPS = sum over pairs of M*L / stage time / 1e9.
Prints one JSON line (see DESIGN.md for every field).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from __graft_entry__ import load_package  # noqa: E402

# Algorithmic VALU work per DP cell ("Roofline"): MSV 1 add + 2 max;
# Viterbi 10 add + 6 max (+1 select for the I state); Forward 17 flops (FMA = 2).
OPS_PER_CELL = {0: 3, 1: 17, 2: 17, 3: 19}
# MI355X vector ALU peak, non-packed f32 lane operations: 256 CU x 4 SIMD x 16 lanes x 2.4 GHz.
VALU_PEAK_TOPS = 256 * 4 * 16 * 2.4e9 / 1e12
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(orc, wl, stage, budget_s=15.0, threads=None):
    """Time the oracle's threaded pair driver on a bounded sample of the workload."""
    threads = threads or min(16, os.cpu_count() or 1)
    hm = [orc.Hmm(name=h.name, m=h.m, mat=h.mat, t=h.t, compo=h.compo) for h in wl.hmms]
    lens = np.array([len(s) for s in wl.seqs], np.int64)
    tot_l = int(lens.sum())
    # grow the profile sample until the estimated time reaches the budget
    rate_guess = {0: 2.0e9, 1: 1.5e8, 2: 1.2e8}[stage] * threads
    npf = 1
    ms = np.array([h.m for h in hm], np.int64)
    while npf < len(hm) and ms[:npf + 1].sum() * tot_l / rate_guess < budget_s:
        npf += 1
    profs = [orc.configure(h) for h in hm[:npf]]
    batch = orc.Batch(profs, wl.seqs)
    fn = {0: batch.msv, 1: batch.viterbi, 2: batch.forward}[stage]
    fn(0, min(64, npf * len(wl.seqs)), threads)  # warm-up
    t0 = time.perf_counter()
    fn(0, npf * len(wl.seqs), threads)
    dt = time.perf_counter() - t0
    cells = int(ms[:npf].sum()) * tot_l
    return {"value": cells / dt / 1e9, "unit": "GCUPS", "cores": threads, "kind": "port",
            "sample": f"{npf} of {len(hm)} profiles x all {len(wl.seqs)} proteins ({cells:.3g} cells, {dt:.1f} s)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", type=int, default=2)
    ap.add_argument("--profile-scale", type=float, default=1.0)
    ap.add_argument("--target-scale", type=float, default=1.0)
    ap.add_argument("--stage", type=int, default=0, help="0 MSV, 1 Viterbi, 2 Forward")
    ap.add_argument("--plain-sample", type=float, default=0.05,
                    help="fraction of the profiles timed in the plain form")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--tune-ms", type=float, default=4000.0, help="autotuner budget (0 = built-in shapes)")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (every box here has one GPU)",
              file=sys.stderr)
        sys.exit(2)

    sxa = load_package()
    from sentinel_x_amd import synth
    t0 = time.time()
    wl = synth.make_workload(args.workload, args.profile_scale, args.target_scale)
    ms = np.array([h.m for h in wl.hmms], np.int64)
    lens = np.array([len(s) for s in wl.seqs], np.int64)
    cells = int(ms.sum()) * int(lens.sum())
    log(f"[bench] workload {wl.name}: {len(ms)} profiles x {len(lens)} proteins, {cells:.3g} cells "
        f"(generated in {time.time() - t0:.1f} s)")
    ctx = sxa.Context(0)
    ctx.set_profiles(wl.hmms)
    ctx.set_targets(wl.seqs)
    stage = args.stage
    if args.tune_ms > 0:
        t0 = time.time()
        ctx.autotune(stage, args.tune_ms)
        log(f"[bench] autotuned shapes (max_m, Q, G, registers) in {time.time() - t0:.1f} s: {ctx.get_shapes(stage)}")

    for i in range(args.warmup):
        _, ms_w = ctx.score_all(stage, sxa.FORM_BATCHED, want=False)
        log(f"[bench] warmup {i}: {ms_w:.1f} ms")
    times = []
    for i in range(args.steps):
        _, ms_s = ctx.score_all(stage, sxa.FORM_BATCHED, want=False)
        times.append(ms_s)
        log(f"[bench] step {i}: {ms_s:.1f} ms, {cells / ms_s / 1e6:.1f} GCUPS")
    ms_step = float(np.median(times))
    gcups = cells / ms_step / 1e6

    # plain form (one launch per profile, one wavefront per pair) on a profile sample
    n_plain = max(1, int(round(len(wl.hmms) * args.plain_sample)))
    ctx_p = sxa.Context(0)
    ctx_p.set_profiles(wl.hmms[:n_plain])
    ctx_p.set_targets(wl.seqs)
    if args.tune_ms > 0:
        ctx_p.set_shapes(stage, ctx.get_shapes(stage))
    plain_cells = int(ms[:n_plain].sum()) * int(lens.sum())
    ctx_p.score_all(stage, sxa.FORM_PLAIN, want=False)
    _, ms_plain = ctx_p.score_all(stage, sxa.FORM_PLAIN, want=False)
    _, ms_pb = ctx_p.score_all(stage, sxa.FORM_BATCHED, want=False)
    plain_gcups = plain_cells / ms_plain / 1e6
    log(f"[bench] plain form on {n_plain} profiles: {ms_plain:.1f} ms = {plain_gcups:.1f} GCUPS "
        f"(batched on the same sample {plain_cells / ms_pb / 1e6:.1f} GCUPS)")
    ctx_p.close()

    achieved = gcups * OPS_PER_CELL[stage] / 1e3  # T lane-ops/s
    roof = {"bound": "valu", "achieved": round(achieved, 3), "peak": round(VALU_PEAK_TOPS, 2), "unit": "TOP/s",
            "frac": round(achieved / VALU_PEAK_TOPS, 4), "traffic": None,
            "ops_per_cell": OPS_PER_CELL[stage]}

    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as orc  # the checker, here timed as the CPU baseline
        cpu = cpu_baseline(orc, wl, stage, args.cpu_budget)
        log(f"[bench] cpu baseline: {cpu}")

    name = sxa.STAGE_NAMES[stage]
    out = {
        "metric": f"GCUPS {name.upper()} batched on 1 MI355X (all pairs)",
        "value": round(gcups, 2), "unit": "GCUPS", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": wl.name, "profiles": len(ms), "proteins": len(lens), "cells": cells,
                   "stage": name, "form": "batched"},
        "plain": {"gcups": round(plain_gcups, 2), "profiles": n_plain, "ratio_batched_over_plain":
                  round(gcups / plain_gcups, 2)},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
