"""Time lane-group shapes of one DP stage on synthetic profiles of one model length.

    python3 sentinel-x_amd/shape_timing.py --stage 0 --m 256 --profiles 100 \
        --shapes 16,16,0 16,16,3 32,8,0 32,8,3

Each shape is q,g,mode (mode = the table mode of sx_set_shapes: 0 LDS, 1 transitions in
registers, 3 / 4 MSV aligned lane blocks with 512 / 1024-thread workgroups), or "auto" for
the built-in choice (the only one for models over 1024 positions).  For each
shape the batched form scores every profile against the targets (log-normal lengths,
median 300) twice; the faster run's device time and rate are printed as one JSON line,
with whether the scores equal those of the first shape (they must: every shape computes
the same bit-exact MSV / Viterbi scores).  Useful to pick a shape table by hand or to
profile one kernel under rocprofv3.
"""
import argparse
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

from __graft_entry__ import load_package  # noqa: E402


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--stage", type=int, default=0, help="0 MSV, 1 Viterbi, 2 Forward, 3 Backward")
    ap.add_argument("--m", type=int, default=256, help="model length")
    ap.add_argument("--profiles", type=int, default=100)
    ap.add_argument("--targets", type=int, default=20000)
    ap.add_argument("--seed", type=int, default=5)
    ap.add_argument("--len", type=int, default=300, help="median target length (log-normal, sigma 0.5)")
    ap.add_argument("--shapes", nargs="+", required=True, help="q,g,mode triples")
    a = ap.parse_args()
    sxa = load_package()
    from sentinel_x_amd import synth
    rng = np.random.default_rng(a.seed)
    hmms = [synth.make_profile(rng, a.m, f"p{i}") for i in range(a.profiles)]
    lens = np.clip(np.exp(rng.normal(np.log(a.len), 0.5, a.targets)), 50, max(2000, 8 * a.len)).astype(int)
    seqs = [synth.background_seq(rng, int(n)) for n in lens]
    cells = a.m * a.profiles * int(lens.sum())
    print(f"# stage {a.stage} M {a.m}: {a.profiles} profiles x {a.targets} targets, {cells:.4g} cells", flush=True)
    ref = None
    with sxa.Context(0) as ctx:
        ctx.set_profiles(hmms)
        ctx.set_targets(seqs)
        for s in a.shapes:
            if s == "auto":
                ctx.set_shapes(a.stage, [])
                q, g, mode = ctx.shape_for(a.stage, a.m)
            else:
                q, g, mode = map(int, s.split(","))
                ctx.set_shapes(a.stage, [(min(q * g, 1024), q, g, mode)])
            out, t1 = ctx.score_all(a.stage, sxa.FORM_BATCHED)
            t2 = ctx.score_all(a.stage, sxa.FORM_BATCHED, want=False)[1]
            if ref is None:
                ref = out
            ms = min(t1, t2)
            print(json.dumps({"stage": a.stage, "m": a.m, "q": q, "g": g, "mode": mode, "ms": round(ms, 3),
                              "gcups": round(cells / ms / 1e6, 1), "same_scores": bool(np.array_equal(out, ref))}),
                  flush=True)


if __name__ == "__main__":
    main()
