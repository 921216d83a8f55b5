"""Throughput of the MI355X Plan7 search path: the MSV -> bias -> Viterbi -> Forward filter
cascade of score_target (sentinel-x src/hmmer/pipeline.rs:159-189) over every
(profile, protein) pair of a synthetic workload, batched (device-resident, graph-replayed)
vs plain (one launch per profile and stage).

A step is one cascade over all pairs.  GCUPS = sum over all pairs of M*L (the MSV cells)
/ step time / 1e9; per-stage GCUPS divide each stage's own cells by its time.  The default
workload is the Pfam-scale set: 20,000 profiles x 50,000 proteins, 6.1e13 MSV cells.
Prints one JSON line; progress goes to stderr.
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

# Algorithmic work of an MSV cell: one f32 add and two f32 max (the max with the B entry
# score, sc + max(prev, xB), and the running row maximum xE).  The kernels issue 2.5
# instructions per cell (a v_max3 folds two cells' xE maxima); packed f32 is not used.
MSV_OPS_PER_CELL = 3
# MI355X's published FP32 vector rate, 157.3 TFLOP/s, counts a fused multiply-add as two
# flops: 78.6 T lane-operations/s (256 CU x 4 SIMD x 32 lanes x 2.4 GHz).  On gfx950
# v_add_f32 issues in about 2.3-2.6 cycles per wave64 instruction but v_max_f32 and
# v_max3_f32 take about 4, so an MSV cell (add + max + half a max3) needs 8 SIMD cycles per
# 64 cells at best: 19.7 T cells/s, 0.75 of this peak in algorithmic ops.
VALU_PEAK_TOPS = 157.3 / 2
# The int16 MSV bound does the same three operations per cell on packed int16 halves: one
# v_pk_add_i16 (saturating) and two v_pk_max_i16 per two cells.  Its peak counts two
# 16-bit lane-operations per 32-bit lane and instruction at the FP32 vector instruction
# rate: 2 x 78.6 = 157.3 T ops/s (numerically the published FP32 rate, which counts an FMA
# as two flops).
VALU16_PEAK_TOPS = 157.3


def measured_traffic(workload, kernel, mode):
    """HBM bytes per launch of the dominant MSV kernel (int16 bound or f32), from the
    rocprofv3 FETCH_SIZE and WRITE_SIZE runs recorded under profiles/ (scaled by the
    copy-kernel calibration given there); None when no record is for this workload and
    kernel."""
    for name in ("msv16_pfam_scale_traffic.json", "msv_pfam_scale_traffic.json"):
        p = os.path.join(ROOT, "profiles", name)
        try:
            with open(p) as f:
                t = json.load(f)
        except (OSError, ValueError):
            continue
        if t.get("workload") != workload:
            continue
        for k in t.get("kernels", [t]):
            if k.get("kernel") == kernel and k.get("table_mode") == mode:
                return int(k["bytes_read"]) + int(k["bytes_written"]), f"profiles/{name}"
    return None, None


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def calibrate(sxa, synth, hmms, n):
    """HMMER-style STATS for every profile, scored by the device kernels themselves."""
    with sxa.Context(0) as cal:
        cal.set_profiles(hmms)

        def scorer(stage, seqs):
            cal.set_targets(seqs)
            out, _ = cal.score_all(stage, sxa.FORM_BATCHED)
            return out

        synth.calibrate(hmms, scorer, n=n)


def host_cpus():
    """The host's logical CPUs (nproc), the CPUs this process may run on, the CPU model,
    and the thread count the CPU legs use: the CPUs this process may use, at most the 16
    host CPUs a one-GPU box allots to its GPU's jobs."""
    n_all = os.cpu_count() or 1
    try:
        n_aff = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        n_aff = n_all
    model = None
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": n_all, "affinity": n_aff, "model": model, "threads": max(1, min(16, n_aff))}


def cpu_baseline(wl, recs, f, budget_s, threads=None):
    """The oracle's threaded cascade on a bounded sample (first k profiles x all proteins);
    with the device's records (recs), the decisions on that sample compared with them."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc  # the checker, timed here as the CPU baseline
    hc = host_cpus()
    threads = threads or hc["threads"]
    lens = np.array([len(s) for s in wl.seqs], np.int64)
    ms = np.array([h.m for h in wl.hmms], np.int64)
    tot_l = int(lens.sum())
    rate_guess = 2.0e9 * threads  # MSV-dominated cells per second
    k = 1
    while k < len(ms) and ms[:k + 1].sum() * tot_l / rate_guess < budget_s:
        k += 1
    hm = [orc.Hmm(name=h.name, m=h.m, mat=h.mat, t=h.t, compo=h.compo, stats=h.stats) for h in wl.hmms[:k]]
    batch = orc.Batch([orc.configure(h) for h in hm], wl.seqs)
    batch.cascade(**f, begin=0, end=min(64, k * len(wl.seqs)), nthreads=threads)  # warm-up
    t0 = time.perf_counter()
    stage, sc = batch.cascade(**f, nthreads=threads)
    dt = time.perf_counter() - t0
    cells = int(ms[:k].sum()) * tot_l
    nt = len(wl.seqs)
    out = {"value": round(cells / dt / 1e9, 3), "unit": "GCUPS", "cores": threads, "kind": "port",
           "sample": f"cascade on {k} of {len(ms)} profiles x all {nt} proteins ({cells:.3g} MSV cells, "
                     f"{dt:.1f} s)",
           "host": hc, "per_core_gcups": round(cells / dt / 1e9 / threads, 3)}
    if recs is None:
        return out
    # decision agreement with the device cascade on the same pairs
    sel = recs[recs["profile"] < k]
    gkey = sel["profile"].astype(np.int64) * nt + sel["target"].astype(np.int64)
    okey = np.flatnonzero(stage >= 1)
    same_set = bool(np.array_equal(np.sort(gkey), okey))
    order = np.argsort(gkey)
    agree = int(np.sum(sel["stage"][order] == stage[np.sort(gkey)])) if same_set else 0
    out["check"] = {"msv_survivor_set_identical": same_set, "stage_agreement": f"{agree}/{len(okey)}"}
    return out


def cpu_forward(wl, budget_s, threads):
    """The oracle's f32 Forward (the reference's dp.rs:229-298) on every protein of the first
    k profiles, within about budget_s: the CPU side of the precision-stress workload, which
    runs Forward on every pair."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc  # the checker, timed here as the CPU baseline
    ms = np.array([h.m for h in wl.hmms], np.int64)
    tot_l = sum(len(x) for x in wl.seqs)
    k = 1
    while k < len(ms) and ms[:k + 1].sum() * tot_l / (0.5e9 * threads) < budget_s:
        k += 1
    hm = [orc.Hmm(name=h.name, m=h.m, mat=h.mat, t=h.t, compo=h.compo, stats=h.stats) for h in wl.hmms[:k]]
    batch = orc.Batch([orc.configure(h) for h in hm], wl.seqs)
    t0 = time.perf_counter()
    batch.forward(nthreads=threads)
    dt = time.perf_counter() - t0
    cells = int(ms[:k].sum()) * tot_l
    return {"value": round(cells / dt / 1e9, 3), "unit": "GCUPS", "cores": threads, "kind": "port",
            "sample": f"f32 Forward on {k} of {len(ms)} profiles x all {len(wl.seqs)} proteins ({cells:.3g} cells, "
                      f"{dt:.1f} s)"}


def plain_leg(sxa, ctx, wl, n_plain, f, graph, bound, vbound):
    """The plain form (one launch per profile and stage) on the first n_plain profiles, and
    the batched form on exactly the same pairs, with the tuned shapes of ctx."""
    key = lambda a: np.lexsort((a["target"], a["profile"]))
    with sxa.Context(0) as cp:
        cp.set_profiles(wl.hmms[:n_plain])
        cp.set_targets(wl.seqs)
        for st in (0, 1, 2, 4, 5):
            cp.set_shapes(st, ctx.get_shapes(st))
        sp_cells = sum(h.m for h in wl.hmms[:n_plain]) * sum(len(s) for s in wl.seqs)
        pl, prec = cp.run_cascade(sxa.FORM_PLAIN, msv_bound=bound, vit_bound=vbound, **f)
        cp.run_cascade(sxa.FORM_BATCHED, use_graph=graph, msv_bound=bound, vit_bound=vbound, **f)
        bt, brec = cp.run_cascade(sxa.FORM_BATCHED, use_graph=graph, msv_bound=bound, vit_bound=vbound, **f)
        same = bool(np.array_equal(prec[key(prec)][["profile", "target", "stage"]],
                                   brec[key(brec)][["profile", "target", "stage"]]))
    return {"profiles": n_plain, "cells": sp_cells, "ms": round(pl.ms_total, 3),
            "gcups": round(sp_cells / pl.ms_total / 1e6, 2),
            "batched_same_pairs_gcups": round(sp_cells / bt.ms_total / 1e6, 2),
            "ratio_batched_over_plain": round(pl.ms_total / bt.ms_total, 2),
            "identical_decisions": same}


SHAPES_DIR = os.path.join(ROOT, "sentinel-x_amd", "shapes")


def install_shapes(ctx, wl, choice, tune_ms, stages):
    """The shape tables of ctx: the committed table for this workload (or the file named
    by choice), else the autotuner's (choice 'tune', or no table for the workload).
    Returns where the tables came from."""
    sys.path.insert(0, os.path.join(ROOT, "sentinel-x_amd"))
    import tune_shapes
    key = tune_shapes.workload_key(wl)
    path = choice if choice not in (None, "tune") else os.path.join(SHAPES_DIR, f"{wl.name}.json")
    if choice != "tune" and os.path.exists(path):
        got = tune_shapes.load(ctx, path)
        if got == key:
            return os.path.relpath(path, ROOT)
        log(f"[bench] {path} is for '{got}', not '{key}': autotuning instead")
    if tune_ms <= 0:
        return "built-in"
    for st in stages:
        t0 = time.time()
        ctx.autotune(st, tune_ms * (3 if st in (0, 4) else 1))  # MSV times 3x larger samples
        log(f"[bench] stage {st} shapes (max_m, Q, G, reg) in {time.time() - t0:.1f} s: {ctx.get_shapes(st)}")
        for row in ctx.autotune_log(st):
            log(f"[tune] stage {st} bucket<= {row[0]:5d} Q={row[1]:3d} G={row[2]:2d} reg={row[3]} "
                f"{row[4]:9.1f} GCUPS")
    return "autotuned"


def kernel_mode(sxa, synth, tune_ms, reps=5, shapes=None):
    """Every pair through every stage (no filters) on the marker-set workload (400 profiles,
    M 80-1500, x 4,000 proteins): batched and plain GCUPS per stage, on identical pairs;
    then the full search (filters on) with its phase times.  Returns the workload, the
    per-stage figures and the batched Forward scores."""
    wl = synth.make_workload(1)
    ms = np.array([h.m for h in wl.hmms], np.int64)
    cells = int(ms.sum()) * sum(len(s) for s in wl.seqs)
    calibrate(sxa, synth, wl.hmms, 200)  # STATS, for the E-values of the Forward-error report
    out = {"workload": f"{wl.name}: {len(ms)} profiles (M {ms.min()}-{ms.max()}) x {len(wl.seqs)} proteins",
           "cells_per_stage": cells}
    with sxa.Context(0) as kc:
        kc.set_profiles(wl.hmms)
        kc.set_targets(wl.seqs)
        out["shape_table"] = install_shapes(kc, wl, shapes, tune_ms, (0, 1, 2, 4))
        for st, name in ((0, "msv"), (1, "viterbi"), (2, "forward")):
            kc.score_all(st, sxa.FORM_BATCHED, want=False)
            tb = float(np.median([kc.score_all(st, sxa.FORM_BATCHED, want=False)[1] for _ in range(reps)]))
            kc.score_all(st, sxa.FORM_PLAIN, want=False)
            tp = float(np.median([kc.score_all(st, sxa.FORM_PLAIN, want=False)[1] for _ in range(3)]))
            out[name] = {"batched_ms": round(tb, 3), "batched_gcups": round(cells / tb / 1e6, 1),
                         "plain_ms": round(tp, 3), "plain_gcups": round(cells / tp / 1e6, 1),
                         "ratio_batched_over_plain": round(tp / tb, 2)}
        fwd, _ = kc.score_all(2, sxa.FORM_BATCHED)
        shapes16 = kc.get_shapes(4)
        # the whole hmmsearch-equivalent search (filters on, E <= 10): where its time goes
        for _ in range(2):
            hits, doms = kc.search(cutoff="evalue", seq_e=10.0, dom_e=10.0, filters=True)
        st = kc.last_search_stats
        # the tail's pairs get Forward, Backward and region finding on the device; what
        # reaches the host is their scores and region lists (bytes_to_host), not the
        # special rows (28 bytes a row a pass: five f32 values and the f64 scale)
        out["search"] = {"hits": int(len(hits)), "domains": int(len(doms)), "pairs_in_tail": int(st["n_pairs"]),
                         "rows_in_tail": int(st["n_rows"]), "regions": int(st["n_regions"]),
                         "envelopes": int(st["n_envelopes"]),
                         "bytes_to_host": int(st["bytes_to_host"]), "special_row_bytes": int(st["n_rows"]) * 56,
                         **{k: round(st[k], 2) for k in ("ms_filters", "ms_specials", "ms_rescore", "ms_tail",
                                                         "ms_report")}}
        # the same search with the envelopes rescored on host threads (the reference's place)
        for _ in range(2):
            hh, dh = kc.search(cutoff="evalue", seq_e=10.0, dom_e=10.0, filters=True, host_rescore=True)
        sh = kc.last_search_stats
        out["search"]["host_rescoring"] = {
            "ms_tail": round(sh["ms_tail"], 2), "threads": os.cpu_count(),
            "same_hits_and_coordinates": bool(
                np.array_equal(hits[["profile", "target", "ndom"]], hh[["profile", "target", "ndom"]]) and
                all(np.array_equal(doms[k], dh[k]) for k in ("hmm_from", "hmm_to", "ali_from", "ali_to",
                                                             "env_from", "env_to")))}
    # the int16 MSV bound covers models up to 1024 positions: timed on those profiles
    short = [h for h in wl.hmms if h.m <= 1024]
    cells16 = sum(h.m for h in short) * sum(len(s) for s in wl.seqs)
    with sxa.Context(0) as kq:
        kq.set_profiles(short)
        kq.set_targets(wl.seqs)
        kq.set_shapes(4, shapes16)
        kq.score_all(4, sxa.FORM_BATCHED, want=False)
        tb = float(np.median([kq.score_all(4, sxa.FORM_BATCHED, want=False)[1] for _ in range(reps)]))
        kq.score_all(4, sxa.FORM_PLAIN, want=False)
        tp = float(np.median([kq.score_all(4, sxa.FORM_PLAIN, want=False)[1] for _ in range(3)]))
        out["msv16_bound"] = {"profiles": len(short), "cells": cells16,
                              "batched_ms": round(tb, 3), "batched_gcups": round(cells16 / tb / 1e6, 1),
                              "plain_ms": round(tp, 3), "plain_gcups": round(cells16 / tp / 1e6, 1),
                              "ratio_batched_over_plain": round(tp / tb, 2)}
    return wl, out, fwd


def precision_gpu(sxa, synth, chunk=500):
    """Forward on the whole precision-stress workload: its 200 profiles (M 1500-2500)
    against all 2,000 proteins (L 8,000-35,000 with low-complexity runs), 400,000 pairs of
    one 64-lane wavefront each, scored chunk proteins at a time (a few seconds a launch).
    The rate is the cells over the launches' summed device time.  Also returns the slice of
    the workload's first 100 proteins and the device scores on it, for the FP64 error check
    and the CPU baseline (the slice's proteins are the same sequences: the generator draws
    them in order)."""
    full = synth.make_workload(3, profile_scale=1.0, target_scale=1.0)
    wl = synth.make_workload(3, profile_scale=1.0, target_scale=0.05)
    n = len(wl.seqs)
    assert all(np.array_equal(a, b) for a, b in zip(wl.seqs, full.seqs[:n]))
    ms = np.array([h.m for h in full.hmms], np.int64)
    cells = int(ms.sum()) * sum(len(s) for s in full.seqs)
    total = 0.0
    with sxa.Context(0) as pc:
        pc.set_profiles(full.hmms)
        pc.set_targets(wl.seqs)
        pc.score_all(2, sxa.FORM_BATCHED, want=False)
        fwd, _ = pc.score_all(2, sxa.FORM_BATCHED)
        for a in range(0, len(full.seqs), chunk):
            pc.set_targets(full.seqs[a:a + chunk])
            _, ms_ = pc.score_all(2, sxa.FORM_BATCHED, want=False)
            total += ms_
    return wl, fwd, {"workload": f"{full.name}: {len(full.hmms)} profiles (M {ms.min()}-{ms.max()}) x "
                                 f"{len(full.seqs)} proteins (L {min(len(x) for x in full.seqs)}-"
                                 f"{max(len(x) for x in full.seqs)})",
                     "cells": cells, "launches": -(-len(full.seqs) // chunk), "forward_ms": round(total, 3),
                     "forward_gcups": round(cells / total / 1e6, 1)}


def single_m200_leg(sxa, synth, f, reps=5):
    """The single-M200 workload (1 profile of 200 positions x 2,000 proteins, 5% homologues):
    the batched and plain cascades and every stage on every pair, GCUPS per stage.  At
    1.4e8 cells a stage it does not fill the device; it is the reference's CPU-sized case."""
    wl = synth.make_workload(0)
    calibrate(sxa, synth, wl.hmms, 200)
    cells = wl.cells
    out = {"workload": f"{wl.name}: 1 profile (M 200) x {len(wl.seqs)} proteins", "cells_per_stage": cells}
    with sxa.Context(0) as c0:
        c0.set_profiles(wl.hmms)
        c0.set_targets(wl.seqs)
        for form, name in ((sxa.FORM_BATCHED, "batched"), (sxa.FORM_PLAIN, "plain")):
            c0.run_cascade(form, **f)
            t = []
            for _ in range(reps):
                st, _ = c0.run_cascade(form, **f)
                t.append(st.ms_total)
            ms = float(np.median(t))
            out[f"cascade_{name}"] = {"ms": round(ms, 3), "gcups": round(cells / ms / 1e6, 1),
                                      "survivors": [int(st.n_msv_pass), int(st.n_bias_pass), int(st.n_vit_pass),
                                                    int(st.n_fwd_pass)]}
        for stage, name in ((0, "msv"), (1, "viterbi"), (2, "forward")):
            c0.score_all(stage, sxa.FORM_BATCHED, want=False)
            ms = float(np.median([c0.score_all(stage, sxa.FORM_BATCHED, want=False)[1] for _ in range(reps)]))
            out[name] = {"ms": round(ms, 3), "gcups": round(cells / ms / 1e6, 1)}
    return wl, out


def mag_sweep_leg(sxa, synth, f, profile_scale, n_genomes, shapes=None, tune_ms=1000.0):
    """The MAG-sweep workload: marker profiles against whole genomes of uneven size.
    packed: every genome's proteins in one target set, so each stage is one batched
    persistent launch per shape over all genomes (graph replay; all genomes complete
    together).  per_genome: the genomes one after another, each its own cascade; the
    latency of a genome is the host wall time from setting its proteins to holding its
    records (p50 / p99 over genomes), and the throughput is the sweep's total cells over
    the summed latencies."""
    t0 = time.time()
    wl = synth.make_workload(4, profile_scale=profile_scale, n_genomes=n_genomes)
    log(f"[bench] mag-sweep workload: {len(wl.hmms)} profiles x {len(wl.seqs)} proteins ({time.time() - t0:.1f} s)")
    calibrate(sxa, synth, wl.hmms, 200)
    ms = np.array([h.m for h in wl.hmms], np.int64)
    lens = np.array([len(x) for x in wl.seqs], np.int64)
    cells = int(ms.sum()) * int(lens.sum())
    out = {"workload": f"{wl.name}: {len(ms)} profiles (M {ms.min()}-{ms.max()}) x {n_genomes} genomes "
                       f"({len(lens)} proteins)", "cells": cells}
    with sxa.Context(0) as cm:
        cm.set_profiles(wl.hmms)
        cm.set_targets(wl.seqs)
        out["shape_table"] = install_shapes(cm, wl, shapes, tune_ms, (4, 0, 1, 2))
        cm.run_cascade(sxa.FORM_BATCHED, use_graph=True, **f)
        t = []
        for _ in range(2):
            st, _ = cm.run_cascade(sxa.FORM_BATCHED, use_graph=True, cap=0, **f)
            t.append(st.ms_total)
        pk = float(min(t))
        out["packed"] = {"ms": round(pk, 3), "gcups": round(cells / pk / 1e6, 1), "hits": int(st.n_fwd_pass)}
        log(f"[bench] mag-sweep packed: {out['packed']}")
        # the whole search over the packed genomes (filters on, E <= 10), envelopes rescored
        # on the device and, for comparison, on host threads
        srch = {}
        for host in (False, True):
            for _ in range(2):
                h, d = cm.search(cutoff="evalue", seq_e=10.0, dom_e=10.0, filters=True, host_rescore=host, **f)
            ss = cm.last_search_stats
            log(f"[bench] mag-sweep search ({'host' if host else 'device'} rescoring): {len(h)} hits")
            srch["host_rescoring" if host else "device_rescoring"] = {
                "hits": int(len(h)), "domains": int(len(d)), "envelopes": int(ss["n_envelopes"]),
                **{k: round(ss[k], 2) for k in ("ms_filters", "ms_specials", "ms_rescore", "ms_tail", "ms_report")}}
            if not host:
                h_dev, d_dev = h, d
        srch["same_hits_and_coordinates"] = bool(
            np.array_equal(h_dev[["profile", "target", "ndom"]], h[["profile", "target", "ndom"]]) and
            all(np.array_equal(d_dev[k], d[k]) for k in ("hmm_from", "hmm_to", "ali_from", "ali_to", "env_from",
                                                         "env_to")))
        out["search"] = srch
        lat = []
        gcells = 0
        for g in range(n_genomes):
            idx = np.flatnonzero(wl.genome_of == g)
            seqs = [wl.seqs[i] for i in idx]
            t0 = time.perf_counter()
            cm.set_targets(seqs)
            st, rec = cm.run_cascade(sxa.FORM_BATCHED, use_graph=False, **f)
            lat.append((time.perf_counter() - t0) * 1e3)
            if g % 25 == 24:
                log(f"[bench] mag-sweep per genome: {g + 1} of {n_genomes}, last {lat[-1]:.1f} ms")
            gcells += int(ms.sum()) * int(sum(len(x) for x in seqs))
        lat = np.array(lat)
        out["per_genome"] = {"latency_ms_p50": round(float(np.percentile(lat, 50)), 2),
                             "latency_ms_p99": round(float(np.percentile(lat, 99)), 2),
                             "latency_ms_max": round(float(lat.max()), 2),
                             "gcups": round(gcells / lat.sum() / 1e6, 1)}
    return wl, out


def ingest_leg(sxa, synth, f, n_profiles=2000, batches=(256, 512, 0)):
    """Model-file ingest: n_profiles Pfam-scale profiles (calibrated) written as an HMMER3
    text file and searched against the 50,000 Pfam-scale proteins (filters on, E <= 10).
    parse_configure: the file parsed and configured into a packed profile cache on one
    host thread and on all of them (sx_convert_hmm_file).  search: the whole job three
    ways, host wall time: the file loaded whole, then searched (sx_load_hmm_file +
    sx_search); streamed a batch of models at a time with parsing overlapped
    (sx_search_hmm_file, for each of `batches`; 0 = the library's default); and streamed
    from the cache at the default batch.  Built-in launch shapes."""
    import shutil
    import tempfile
    wl = synth.make_workload(2, profile_scale=n_profiles / 20000.0)
    calibrate(sxa, synth, wl.hmms, 200)
    d = tempfile.mkdtemp(prefix="sx_ingest_")
    try:
        path, cache = os.path.join(d, "lib.hmm"), os.path.join(d, "lib.sxpc")
        synth.write_hmm(path, wl.hmms)
        threads = host_cpus()["threads"]
        ms = np.array([h.m for h in wl.hmms])
        out = {"workload": f"{len(wl.hmms)} pfam-scale profiles (M {ms.min()}-{ms.max()}) x {len(wl.seqs)} proteins",
               "file_mb": round(os.path.getsize(path) / 2**20, 1), "threads": threads}
        for th in (1, threads):
            t0 = time.perf_counter()
            sxa.convert_hmm_file(path, cache, th)
            dt = time.perf_counter() - t0
            out[f"parse_configure_{th}t"] = {"s": round(dt, 3), "profiles_per_s": round(len(wl.hmms) / dt, 1)}
        out["cache_mb"] = round(os.path.getsize(cache) / 2**20, 1)
        kw = dict(cutoff="evalue", seq_e=10.0, dom_e=10.0, filters=True, **f)
        with sxa.Context(0) as c:
            c.set_targets(wl.seqs)
            c.load_hmm_file(cache)
            c.search(**kw)  # warm-up: kernels loaded, buffers sized
            c.set_profiles([])
            t0 = time.perf_counter()
            c.load_hmm_file(path)
            t1 = time.perf_counter()
            h0, d0 = c.search(**kw)
            t2 = time.perf_counter()
            out["load_then_search"] = {"s": round(t2 - t0, 3), "load_s": round(t1 - t0, 3), "hits": int(len(h0))}
            runs = [(f"streamed_batch_{b or 'default'}", path, b) for b in batches]
            runs.append(("streamed_from_cache", cache, 0))
            for name, src, b in runs:
                t0 = time.perf_counter()
                h1, d1, _ = c.search_hmm_file(src, batch=b, **kw)
                dt = time.perf_counter() - t0
                st = c.last_search_stats
                out[name] = {"s": round(dt, 3), "ingest_ms_not_overlapped": round(st["ms_ingest"], 1),
                             "filters_ms": round(st["ms_filters"], 1),
                             "same_hits": bool(np.array_equal(h0, h1) and np.array_equal(d0, d1))}
        return out
    finally:
        shutil.rmtree(d, ignore_errors=True)


def inclusion_parity(sxa, wl, f, threads):
    """The whole search (filters on) on the device against the reference's path on the
    CPU, on every pair of the workload: the oracle's cascade over all pairs, its Forward and
    Backward special rows for the pairs it passes, and the host tail (sx_search_host) from
    those rows.  Compared per cutoff (E <= 10, the models' GA, NC and TC): the hits, their
    domains' coordinates and reporting flags, and the scores (search.rs:131-210 decides
    inclusion from them)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc  # the checker
    hm = [orc.Hmm(name=h.name, m=h.m, mat=h.mat, t=h.t, compo=h.compo, stats=h.stats, ga=h.ga, tc=h.tc, nc=h.nc)
          for h in wl.hmms]
    profs = [orc.configure(h) for h in hm]
    t0 = time.perf_counter()
    stage_o, _ = orc.Batch(profs, wl.seqs).cascade(**f, nthreads=threads)
    idx = np.flatnonzero(stage_o == 4)
    pp, tt = (idx // len(wl.seqs)).astype(np.uint32), (idx % len(wl.seqs)).astype(np.uint32)
    fwd, bwd, fsp, bsp = [], [], [], []
    for p, t in zip(pp, tt):
        v, sp = orc.forward(profs[p], wl.seqs[t], specials=True)
        b, spb = orc.backward(profs[p], wl.seqs[t])
        fwd.append(v)
        fsp.append(sp)
        bwd.append(float(b))
        bsp.append(spb)
    out = {"pairs": len(profs) * len(wl.seqs), "pairs_in_tail_cpu": int(len(idx)),
           "oracle_s": round(time.perf_counter() - t0, 2)}
    coords = ("hmm_from", "hmm_to", "ali_from", "ali_to", "env_from", "env_to", "reported")
    with sxa.Context(0) as c:
        c.set_profiles(wl.hmms)
        c.set_targets(wl.seqs)
        for cut in ("evalue", "ga", "nc", "tc"):
            kw = dict(cutoff=cut, seq_e=10.0, dom_e=10.0, filters=True, **f)
            h, d = c.search(**kw)
            hh, dh = sxa.search_host(hm, wl.seqs, pp, tt, fwd, bwd, fsp, bsp, **kw)
            same = bool(np.array_equal(h[["profile", "target", "ndom"]], hh[["profile", "target", "ndom"]]) and
                        all(np.array_equal(d[k], dh[k]) for k in coords))
            out[cut] = {"hits": int(len(h)), "domains": int(len(d)), "identical_hits_domains_and_flags": same,
                        "max_score_diff_bits": float(np.abs(h["score"] - hh["score"]).max()) if same and len(h)
                        else None}
    return out


def forward_error(wl, fwd_gpu, k, threads, z=None):
    """Device Forward (f32 score, as the reference returns it) against the oracle's FP64
    Forward on the first k profiles x all proteins: nats error, and the inclusion decisions
    each score gives on the Forward bit score (Forward minus the null score, in bits) at the
    models' GA / TC / NC sequence cutoffs and at E <= 10 (E from the model's Forward STATS
    line, Z = the number of proteins).  Decisions after null2 and domain definition are not
    part of this library."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc  # the checker
    hm = [orc.Hmm(name=h.name, m=h.m, mat=h.mat, t=h.t, compo=h.compo, stats=h.stats) for h in wl.hmms[:k]]
    t0 = time.perf_counter()
    ref = orc.Batch([orc.configure(h) for h in hm], wl.seqs).forward64(nthreads=threads).reshape(k, len(wl.seqs))
    dt = time.perf_counter() - t0
    gpu = fwd_gpu[:k].astype(np.float32).astype(np.float64)
    fin = np.isfinite(ref)
    err = np.abs(gpu[fin] - ref[fin])
    out = {"pairs": int(fin.sum()), "max_nats": float(err.max()), "median_nats": float(np.median(err)),
           "p99_nats": float(np.quantile(err, 0.99)), "within_1e-3": bool(err.max() <= 1e-3),
           "same_finiteness": bool(np.array_equal(np.isfinite(gpu), fin)), "oracle_s": round(dt, 2)}
    nulls = np.array([float(orc.null_score(len(x))) for x in wl.seqs])
    bits_g = (gpu - nulls[None, :]) / math.log(2.0)
    bits_r = (ref - nulls[None, :]) / math.log(2.0)
    dis = {}
    for cut in ("ga", "tc", "nc"):
        c = np.array([getattr(h, cut)[0] for h in wl.hmms[:k]])[:, None]
        dis[cut] = int(np.sum((bits_g >= c) != (bits_r >= c)))
    if all("FORWARD" in h.stats for h in wl.hmms[:k]):
        zz = z or len(wl.seqs)
        disagree = 0
        for p in range(k):
            tau, lam = wl.hmms[p].stats["FORWARD"]
            for t in np.flatnonzero(fin[p]):
                eg = zz * math.exp(orc.exp_logsurv(float(bits_g[p, t]), tau, lam))
                er = zz * math.exp(orc.exp_logsurv(float(bits_r[p, t]), tau, lam))
                disagree += int((eg <= 10.0) != (er <= 10.0))
        dis["evalue_10"] = disagree
    out["decision_disagreements"] = dis
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", type=int, default=2)
    ap.add_argument("--profile-scale", type=float, default=1.0)
    ap.add_argument("--target-scale", type=float, default=1.0)
    ap.add_argument("--plain-sample", type=float, default=0.02, help="fraction of profiles run in the plain form (0 = skip)")
    ap.add_argument("--shapes", default=None,
                    help="shape table file (default sentinel-x_amd/shapes/<workload>.json when it exists and "
                         "matches the workload); 'tune' to autotune instead")
    ap.add_argument("--tune-ms", type=float, default=3000.0,
                    help="autotuner budget per stage when tuning, three times this for MSV (0 = built-in shapes)")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--no-bound", action="store_true",
                    help="f32 MSV on every pair (no int16 MSV bound in front of it)")
    ap.add_argument("--vit-bound", action="store_true",
                    help="run the int16 Viterbi bound in front of the f32 Viterbi in the timed cascade")
    ap.add_argument("--no-f32-leg", action="store_true",
                    help="skip timing the cascade without the bound beside the headline run")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--f1", type=float, default=0.02)
    ap.add_argument("--f2", type=float, default=1e-3)
    ap.add_argument("--f3", type=float, default=1e-5)
    ap.add_argument("--no-kernel-mode", action="store_true",
                    help="skip the every-pair-every-stage runs and the Forward-error report")
    ap.add_argument("--fwd-err-profiles", type=int, default=40,
                    help="marker-set profiles (x all proteins) checked against the FP64 Forward")
    ap.add_argument("--no-single", action="store_true", help="skip the single-M200 leg")
    ap.add_argument("--ingest", type=int, default=2000,
                    help="profiles of the model-file ingest leg (0 = skip)")
    ap.add_argument("--mag-sweep", default="slice", choices=("slice", "full", "none"),
                    help="MAG-sweep leg: 1,000 markers x 40 genomes (slice), 5,000 x 200 (full), or none")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (every box here has one GPU)",
              file=sys.stderr)
        sys.exit(2)
    f = dict(f1=args.f1, f2=args.f2, f3=args.f3)

    sxa = load_package()
    from sentinel_x_amd import synth
    t0 = time.time()
    wl = synth.make_workload(args.workload, args.profile_scale, args.target_scale)
    ms = np.array([h.m for h in wl.hmms], np.int64)
    lens = np.array([len(s) for s in wl.seqs], np.int64)
    cells = int(ms.sum()) * int(lens.sum())
    log(f"[bench] workload {wl.name}: {len(ms)} profiles x {len(lens)} proteins, {cells:.3g} MSV cells "
        f"(generated in {time.time() - t0:.1f} s)")
    t0 = time.time()
    calibrate(sxa, synth, wl.hmms, 200)
    log(f"[bench] calibrated STATS on the device in {time.time() - t0:.1f} s")

    ctx = sxa.Context(0)
    ctx.set_profiles(wl.hmms)
    ctx.set_targets(wl.seqs)
    bound = not args.no_bound
    vbound = bool(args.vit_bound)
    shape_src = install_shapes(ctx, wl, args.shapes, args.tune_ms, (4, 0, 5, 1, 2) if bound else (0, 5, 1, 2))
    log(f"[bench] shape tables: {shape_src}")
    graph = not args.no_graph
    for i in range(args.warmup):
        s, _ = ctx.run_cascade(sxa.FORM_BATCHED, use_graph=graph, msv_bound=bound, vit_bound=vbound, **f)
        log(f"[bench] warmup {i}: {s.ms_total:.1f} ms")
    times = []
    for i in range(args.steps):
        s, recs = ctx.run_cascade(sxa.FORM_BATCHED, use_graph=graph, msv_bound=bound, vit_bound=vbound, **f)
        times.append(s.ms_total)
        log(f"[bench] step {i}: {s.ms_total:.1f} ms = {cells / s.ms_total / 1e6:.1f} GCUPS; survivors "
            f"MSV {s.n_msv_pass} bias {s.n_bias_pass} Viterbi {s.n_vit_pass} Forward {s.n_fwd_pass}; "
            f"bound candidates MSV {s.n_msv_cand} Viterbi {s.n_vit_cand}")
    ms_step = float(np.median(times))
    gcups = cells / ms_step / 1e6
    # the same cascade with the f32 MSV and Viterbi on every pair of their stages (no
    # bounds): its time, and the same records (decisions and scores; a record the Viterbi
    # bound rejected carries no f32 Viterbi score)
    f32_leg = None
    if bound and not args.no_f32_leg:
        ctx.run_cascade(sxa.FORM_BATCHED, use_graph=graph, msv_bound=False, vit_bound=False, **f)
        s0, recs0 = ctx.run_cascade(sxa.FORM_BATCHED, use_graph=graph, msv_bound=False, vit_bound=False, **f)
        key0 = lambda a: np.lexsort((a["target"], a["profile"]))
        a, b = recs0[key0(recs0)], recs[key0(recs)]
        scored = ~np.isnan(b["vit"])
        same = bool(np.array_equal(a[["profile", "target", "stage"]], b[["profile", "target", "stage"]])
                    and all(np.array_equal(a[k], b[k], equal_nan=True) for k in ("msv", "bias", "fwd64"))
                    and np.array_equal(a["vit"][scored], b["vit"][scored])
                    and bool(np.all(b["stage"][~scored & (a["stage"] >= 2)] == 2)))
        f32_leg = {"ms_per_step": round(s0.ms_total, 3), "gcups": round(cells / s0.ms_total / 1e6, 2),
                   "speedup_of_bounds": round(s0.ms_total / ms_step, 3), "same_decisions_and_scores": same}
        log(f"[bench] f32 cascade (no bounds): {f32_leg}")
    # per-stage split from one eager (event-timed) run of the same chain
    e, erecs = ctx.run_cascade(sxa.FORM_BATCHED, use_graph=False, msv_bound=bound, vit_bound=vbound, **f)
    ls = ctx.cascade_launches()  # this run's launches (the legs below run the cascade again)
    stages = {
        "msv": {"ms": round(e.ms_msv, 3), "cells": int(e.cells_msv), "gcups": round(e.cells_msv / e.ms_msv / 1e6, 1)},
        "bias+compaction": {"ms": round(e.ms_bias, 3), "pairs": int(e.n_msv_pass),
                            "host_recheck": {"pairs": int(e.n_bias_recheck), "ms": round(e.ms_recheck, 3)}},
        "viterbi": {"ms": round(e.ms_vit, 3), "cells": int(e.cells_vit),
                    "gcups": round(e.cells_vit / max(e.ms_vit, 1e-9) / 1e6, 1)},
        "forward": {"ms": round(e.ms_fwd, 3), "cells": int(e.cells_fwd),
                    "gcups": round(e.cells_fwd / max(e.ms_fwd, 1e-9) / 1e6, 1)},
    }
    if bound:
        # the MSV stage = the bound over every pair + candidate compaction + f32 MSV of the
        # candidates; its time includes all three
        stages["msv"]["bound"] = {"ms": round(e.ms_msv_bound, 3),
                                  "gcups": round(e.cells_msv / max(e.ms_msv_bound, 1e-9) / 1e6, 1)}
        stages["msv"]["candidates"] = {"pairs": int(e.n_msv_cand), "cells": int(e.cells_msv_exact),
                                       "ms": round(e.ms_msv - e.ms_msv_bound, 3),
                                       "fraction_of_pairs": round(e.n_msv_cand / max(e.n_pairs, 1), 5)}
    # the int16 Viterbi bound in front of the f32 Viterbi (one eager run, or the timed
    # cascade's own split when it runs with the bound)
    ev = e
    if not vbound:
        ev, _ = ctx.run_cascade(sxa.FORM_BATCHED, use_graph=False, msv_bound=bound, vit_bound=True, **f)
    viterbi_bound = {"in_timed_cascade": vbound, "viterbi_stage_ms": round(ev.ms_vit, 3),
                     "bound_ms": round(ev.ms_vit_bound, 3),
                     "bound_gcups": round(ev.cells_vit / max(ev.ms_vit_bound, 1e-9) / 1e6, 1),
                     "candidates": int(ev.n_vit_cand),
                     "fraction_of_pairs": round(ev.n_vit_cand / max(ev.n_bias_pass, 1), 5),
                     "f32_viterbi_stage_ms_without_bound": None if vbound else round(e.ms_vit, 3)}
    log(f"[bench] Viterbi bound: {viterbi_bound}")
    log(f"[bench] eager split: {stages}")
    key = lambda a: np.lexsort((a["target"], a["profile"]))
    same_as_eager = bool(np.array_equal(recs[key(recs)][["profile", "target", "stage"]],
                                        erecs[key(erecs)][["profile", "target", "stage"]]))

    # plain form on a profile sample, and the batched form on exactly the same pairs
    plain = None
    if args.plain_sample > 0:
        plain = plain_leg(sxa, ctx, wl, max(1, int(round(len(ms) * args.plain_sample))), f, graph, bound, vbound)
        log(f"[bench] plain form: {plain}")

    # dominant kernel: the longest MSV launch of the eager run (HIP events on the launch
    # stream): an int16 bound launch over all pairs, or an f32 MSV launch without the bound
    mstage = 4 if bound else 0
    msv = ls[ls["stage"] == mstage]
    d = msv[int(np.argmax(msv["ms"]))]
    peak = VALU16_PEAK_TOPS if bound else VALU_PEAK_TOPS
    achieved = float(d["cells"]) * MSV_OPS_PER_CELL / (float(d["ms"]) * 1e-3) / 1e12
    ms_launches = float(msv["ms"].sum())
    stage_achieved = e.cells_msv * MSV_OPS_PER_CELL / (e.ms_msv * 1e-3) / 1e12
    kname = f"k_tiles<Q={int(d['q'])},G={int(d['g'])},{'MSV16' if bound else 'MSV'}>"
    traffic, traffic_src = measured_traffic(f"{wl.name}: {len(ms)} profiles x {len(lens)} proteins", kname,
                                            int(d["hoist"]))
    roof = {"bound": "valu", "achieved": round(achieved, 3), "peak": round(peak, 2), "unit": "TOP/s",
            "frac": round(achieved / peak, 4), "traffic": traffic, "traffic_source": traffic_src,
            "kernel": kname, "table_mode": int(d["hoist"]), "launch_ms": round(float(d["ms"]), 3),
            "cells": int(d["cells"]), "ops_per_cell": MSV_OPS_PER_CELL,
            "all_launches_of_kernel_stage": {"launches": int(len(msv)), "ms": round(ms_launches, 3),
                                             "achieved": round(e.cells_msv * MSV_OPS_PER_CELL /
                                                               (ms_launches * 1e-3) / 1e12, 3)},
            "msv_stage": {"achieved": round(stage_achieved, 3), "frac": round(stage_achieved / peak, 4)}}
    launches = [{"stage": int(r["stage"]), "q": int(r["q"]), "g": int(r["g"]), "reg": int(r["hoist"]),
                 "ms": round(float(r["ms"]), 3), "cells": int(r["cells"]),
                 "gcups": round(float(r["cells"]) / max(float(r["ms"]), 1e-9) / 1e6, 1)}
                for r in ls]
    log(f"[bench] launches: {launches}")

    kmode = prec = wl1 = wl3 = None
    if not args.no_kernel_mode:
        t0 = time.time()
        wl1, kmode, fwd1 = kernel_mode(sxa, synth, min(args.tune_ms, 1000.0) if args.tune_ms > 0 else 0,
                                       shapes="tune" if args.shapes == "tune" else None)
        log(f"[bench] kernel mode ({time.time() - t0:.1f} s): {kmode}")
        t0 = time.time()
        wl3, fwd3, prec = precision_gpu(sxa, synth)
        log(f"[bench] precision-stress Forward ({time.time() - t0:.1f} s): {prec}")

    single = mag = wl0 = wl4 = None
    if not args.no_single:
        t0 = time.time()
        wl0, single = single_m200_leg(sxa, synth, f)
        log(f"[bench] single-M200 ({time.time() - t0:.1f} s): {single}")
    if args.mag_sweep != "none":
        t0 = time.time()
        ps, ng = (1.0, 200) if args.mag_sweep == "full" else (0.2, 40)
        wl4, mag = mag_sweep_leg(sxa, synth, f, ps, ng, shapes="tune" if args.shapes == "tune" else None,
                            tune_ms=min(args.tune_ms, 1000.0) if args.tune_ms > 0 else 0)
        log(f"[bench] mag-sweep ({time.time() - t0:.1f} s): {mag}")

    ingest = None
    if args.ingest > 0:
        t0 = time.time()
        ingest = ingest_leg(sxa, synth, f, n_profiles=args.ingest)
        log(f"[bench] ingest ({time.time() - t0:.1f} s): {ingest}")

    cpu = fwd_err = None
    if not args.no_cpu:
        cpu = cpu_baseline(wl, recs, f, args.cpu_budget)
        log(f"[bench] cpu baseline: {cpu}")
        if kmode is not None:
            threads = host_cpus()["threads"]
            k1, k3 = min(args.fwd_err_profiles, len(wl1.hmms)), min(4, len(wl3.hmms))
            fwd_err = {"marker_set": forward_error(wl1, fwd1, k1, threads),
                       "precision": forward_error(wl3, fwd3, k3, threads)}
            fwd_err["marker_set"]["sample"] = f"first {k1} profiles x all {len(wl1.seqs)} proteins"
            fwd_err["precision"]["sample"] = f"first {k3} profiles x the workload's first {len(wl3.seqs)} proteins"
            log(f"[bench] forward error vs FP64: {fwd_err}")
            t0 = time.time()
            fwd_err["inclusion_parity_marker_set"] = inclusion_parity(sxa, wl1, f, threads)
            log(f"[bench] inclusion parity ({time.time() - t0:.1f} s): {fwd_err['inclusion_parity_marker_set']}")
        # the CPU side of the other workloads, each on a bounded sample
        threads = host_cpus()["threads"]
        cpu_cfg = {}
        if wl0 is not None:
            cpu_cfg["single_m200"] = cpu_baseline(wl0, None, f, 2.0)
        if wl1 is not None:
            cpu_cfg["marker_set"] = cpu_baseline(wl1, None, f, 3.0)
            cpu_cfg["precision_forward"] = cpu_forward(wl3, 4.0, threads)
        if wl4 is not None:
            cpu_cfg["mag_sweep"] = cpu_baseline(wl4, None, f, 3.0)
        cpu["per_workload"] = cpu_cfg
        log(f"[bench] cpu per workload: {cpu_cfg}")

    out = {
        "metric": "GCUPS of the MSV->bias->Viterbi->Forward cascade over all pairs, batched, 1 MI355X",
        "value": round(gcups, 2), "unit": "GCUPS", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "i16+f32" if bound else "f32", "data": "synthetic",
        "config": {"workload": f"{wl.name}: {len(ms)} profiles x {len(lens)} proteins", "cells_msv": cells,
                   "f1": args.f1, "f2": args.f2, "f3": args.f3, "graph": graph, "msv_bound": bound,
                   "viterbi_bound": vbound, "shape_table": shape_src},
        "f32_cascade_without_bounds": f32_leg,
        "viterbi_bound": viterbi_bound,
        "stages": stages,
        "survivors": {"msv": int(e.n_msv_pass), "bias": int(e.n_bias_pass), "viterbi": int(e.n_vit_pass),
                      "forward": int(e.n_fwd_pass)},
        "graph_matches_eager": same_as_eager,
        "launches": launches,
        "plain": plain,
        "roofline": roof,
        "cpu_baseline": cpu,
        "kernel_mode": kmode,
        "precision_forward": prec,
        "forward_error_vs_fp64": fwd_err,
        "single_m200": single,
        "mag_sweep": mag,
        "ingest": ingest,
    }
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
