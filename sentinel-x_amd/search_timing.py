"""Time the phases of a whole search (sx_search, filters on, E <= 10) on a synthetic workload,
with the envelopes rescored on the device and on host threads.

    python3 sentinel-x_amd/search_timing.py --workload mag --profile-scale 0.2 --genomes 40

Workloads: "marker" (400 profiles x 4,000 proteins, 3% homologues) or "mag" (marker
profiles x whole genomes, 2% homologues).  Models are calibrated (STATS) first.  Each
form runs twice; the second run's phase times (sx_results_stats) are printed as one JSON
line per form, with whether the two forms report the same hits and coordinates.  Useful
to profile the envelope rescoring kernels (sxe_rescore*) under rocprofv3.
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
    ap.add_argument("--workload", choices=("marker", "mag"), default="mag")
    ap.add_argument("--profile-scale", type=float, default=0.2)
    ap.add_argument("--genomes", type=int, default=40)
    ap.add_argument("--shapes", default=None,
                    help="shape table file to install (stages it lists); default: the built-in shapes")
    a = ap.parse_args()
    sxa = load_package()
    from sentinel_x_amd import synth
    import bench
    if a.workload == "marker":
        wl = synth.make_workload(1)
    else:
        wl = synth.make_workload(4, profile_scale=a.profile_scale, n_genomes=a.genomes)
    bench.calibrate(sxa, synth, wl.hmms, 200)
    print(f"# {wl.name}: {len(wl.hmms)} profiles x {len(wl.seqs)} proteins", flush=True)
    res = {}
    with sxa.Context(0) as c:
        c.set_profiles(wl.hmms)
        c.set_targets(wl.seqs)
        if a.shapes:
            with open(a.shapes) as fh:
                for st, v in json.load(fh)["stages"].items():
                    c.set_shapes(int(st), [tuple(r) for r in v["table"]])
        for host in (False, True):
            for _ in range(2):
                h, d = c.search(cutoff="evalue", seq_e=10.0, dom_e=10.0, filters=True, host_rescore=host)
            st = c.last_search_stats
            res[host] = (h, d)
            print(json.dumps({"rescoring": "host" if host else "device", "hits": int(len(h)), "domains": int(len(d)),
                              **{k: (round(v, 2) if isinstance(v, float) else v) for k, v in st.items()}}),
                  flush=True)
    (h0, d0), (h1, d1) = res[False], res[True]
    same = bool(np.array_equal(h0[["profile", "target", "ndom"]], h1[["profile", "target", "ndom"]]) and
                all(np.array_equal(d0[k], d1[k]) for k in ("hmm_from", "hmm_to", "ali_from", "ali_to",
                                                           "env_from", "env_to", "reported")))
    print(json.dumps({"same_hits_and_coordinates": same}), flush=True)


if __name__ == "__main__":
    main()
