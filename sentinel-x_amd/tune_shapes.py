"""Autotune the lane-group shape tables of every cascade stage on a benchmark workload and
write them to a JSON file that bench.py (and any caller, through sx_set_shapes) loads.

    python3 sentinel-x_amd/tune_shapes.py --workload 2 --out sentinel-x_amd/shapes/pfam-scale.json

The tables are a property of the kernels and the workload's length distributions, not of
the box; loading a committed table makes the launches, and so the dominant kernel that
the committed profiles describe, the same on every run.  The file records the workload,
the per-stage budgets, each stage's table (rows max_m, q, g, table mode) and the
autotuner's measurements.
"""
import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, ROOT)

from __graft_entry__ import load_package  # noqa: E402

# stage order and budget factor: the int16 MSV bound first (it decides the candidates), MSV
# with three times the budget (its samples are larger), then the Viterbi bound, Viterbi and
# Forward
STAGES = ((4, 3), (0, 3), (5, 1), (1, 1), (2, 1))


def workload_key(wl):
    return f"{wl.name}: {len(wl.hmms)} profiles x {len(wl.seqs)} proteins"


def tune(sxa, ctx, tune_ms, log=print, only=None):
    """Autotune every stage (or the stages in `only`) on ctx's profiles and targets; returns
    {stage: {table, log}}."""
    out = {}
    for st, f in STAGES:
        if only is not None and st not in only:
            continue
        t0 = time.time()
        ctx.autotune(st, tune_ms * f)
        out[str(st)] = {"table": [list(r) for r in ctx.get_shapes(st)],
                        "log": [list(r) for r in ctx.autotune_log(st)]}
        log(f"[tune] stage {st}: {len(out[str(st)]['table'])} rows in {time.time() - t0:.1f} s")
    return out


def load(ctx, path):
    """Install the tables of a file written by this tool; returns its workload key."""
    with open(path) as fh:
        d = json.load(fh)
    for st, v in d["stages"].items():
        ctx.set_shapes(int(st), [tuple(r) for r in v["table"]])
    return d["workload"]


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--workload", type=int, default=2)
    ap.add_argument("--profile-scale", type=float, default=1.0)
    ap.add_argument("--target-scale", type=float, default=1.0)
    ap.add_argument("--tune-ms", type=float, default=3000.0, help="per-stage budget (x3 for MSV and the bound)")
    ap.add_argument("--out", required=True)
    ap.add_argument("--stages", type=int, nargs="+", default=None,
                    help="tune only these stages and keep the others' tables from --out (which must exist)")
    a = ap.parse_args()
    prev = None
    if a.stages:
        with open(a.out) as fh:
            prev = json.load(fh)
    sxa = load_package()
    from sentinel_x_amd import synth
    import bench
    wl = synth.make_workload(a.workload, a.profile_scale, a.target_scale)
    bench.calibrate(sxa, synth, wl.hmms, 200)
    with sxa.Context(0) as ctx:
        ctx.set_profiles(wl.hmms)
        ctx.set_targets(wl.seqs)
        stages = tune(sxa, ctx, a.tune_ms, log=lambda *m: print(*m, file=sys.stderr, flush=True), only=a.stages)
    out = {"workload": workload_key(wl), "tune_ms": a.tune_ms, "stages": stages}
    if prev is not None:
        if prev["workload"] != out["workload"]:
            raise SystemExit(f"{a.out} holds tables of another workload ({prev['workload']})")
        merged = dict(prev["stages"])
        merged.update(stages)
        out["stages"] = {k: merged[k] for k in sorted(merged, key=lambda k: [s for s, _ in STAGES].index(int(k)))}
        out["tune_ms"] = prev["tune_ms"]
        out["tune_ms_by_stage"] = {**prev.get("tune_ms_by_stage", {}), **{str(st): a.tune_ms for st in a.stages}}
    os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
    with open(a.out, "w") as fh:
        json.dump(out, fh, indent=1)
    print(f"wrote {a.out}", flush=True)


if __name__ == "__main__":
    main()
