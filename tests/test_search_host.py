"""CPU checks of the host part of the search (sx_search_host): the rest of score_target
(pipeline.rs:186-294) and `report` (search.rs:160-210), fed with the oracle's Forward /
Backward special rows.

The end-to-end fixture is HMMER 3.4 `hmmsearch --max --domtblout` output for two
simulated families (tests/golden/hmmer34_expected.tsv, 20 rows), checked as the
reference's own test does (search.rs:316-355): sequence score, sequence bias and domain
score within 0.1 bit, hmm / ali / env coordinates exact.  The tiny-model cases restate
pipeline.rs:597-616 and search.rs:285-306."""
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")


def read_faa(path):
    names, seqs = [], []
    for line in open(path):
        line = line.strip()
        if line.startswith(">"):
            names.append(line[1:].strip())
            seqs.append(bytearray())
        elif names:
            seqs[-1].extend(line.encode())
    return names, [bytes(s) for s in seqs]


def oracle_inputs(orc, hmms, seqs):
    """All pairs with the oracle's Forward / Backward scores and special rows."""
    profs = [orc.configure(h) for h in hmms]
    prof, tgt, fwd, bwd, fsp, bsp = [], [], [], [], [], []
    for p, pr in enumerate(profs):
        for t, d in enumerate(seqs):
            f, fs = orc.forward(pr, d, specials=True)
            b, bs = orc.backward(pr, d)
            prof.append(p)
            tgt.append(t)
            fwd.append(f)
            bwd.append(float(b))
            fsp.append(fs)
            bsp.append(bs)
    return prof, tgt, fwd, bwd, fsp, bsp


def check_hmmer34(hits, doms, pnames, tnames):
    rows = [r.split("\t") for r in open(os.path.join(GOLDEN, "hmmer34_expected.tsv")).read().splitlines() if r]
    assert len(rows) == 20
    for f in rows:
        tgt, q = f[0], f[1]
        sel = [h for h in hits if tnames[h["target"]] == tgt and pnames[h["profile"]] == q]
        assert len(sel) == 1, f"missing hit {tgt} {q}"
        h = sel[0]
        assert abs(h["score"] - float(f[2])) <= 0.1, (tgt, q, h["score"], f[2])
        assert abs(h["bias"] - float(f[3])) <= 0.1, (tgt, q, h["bias"], f[3])
        d = doms[h["dom_first"]:h["dom_first"] + h["ndom"]]
        rep = d[d["reported"] != 0]
        dd = rep[int(f[4]) - 1]
        assert abs(dd["score"] - float(f[5])) <= 0.1, (tgt, q, f[4], dd["score"], f[5])
        coords = [int(dd[k]) for k in ("hmm_from", "hmm_to", "ali_from", "ali_to", "env_from", "env_to")]
        assert coords == [int(x) for x in f[7:13]], (tgt, q, f[4], coords, f[7:13])


def test_hmmer34_fixture_scores_and_coordinates(orc, sxa):
    hmms = orc.parse_hmm_file(os.path.join(GOLDEN, "hmmer34_pair.hmm"))
    tnames, raw = read_faa(os.path.join(GOLDEN, "hmmer34_targets.faa"))
    seqs = [orc.encode(s) for s in raw]
    prof, tgt, fwd, bwd, fsp, bsp = oracle_inputs(orc, hmms, seqs)
    hits, doms = sxa.search_host(hmms, seqs, prof, tgt, fwd, bwd, fsp, bsp, cutoff="evalue", seq_e=10.0,
                                 dom_e=10.0)
    # ordered by profile, then target (search.rs:225-233)
    keys = list(zip(hits["profile"], hits["target"]))
    assert keys == sorted(keys)
    check_hmmer34(hits, doms, [h.name for h in hmms], tnames)


def _tiny(orc):
    return orc.parse_hmm_file(os.path.join(GOLDEN, "tiny.hmm"))[0]


def test_tiny_finds_both_copies_of_a_motif(orc, sxa):
    h = _tiny(orc)
    seqs = [orc.encode(b"GGGGGMKAGGGGGGGGGGGGGGGGGGGGMKAGGGGG")]
    ins = oracle_inputs(orc, [h], seqs)
    hits, doms = sxa.search_host([h], seqs, *ins, cutoff="bits", seq_t=-1e9, dom_t=-1e9)
    assert len(hits) == 1 and hits[0]["ndom"] == 2
    assert hits[0]["score"] > 0.0
    assert (doms[0]["ali_from"], doms[0]["ali_to"]) == (6, 8)


def test_tiny_bits_threshold_reports_motif_carrier(orc, sxa):
    h = _tiny(orc)
    seqs = [orc.encode(b"GGGGMKAGGGG"), orc.encode(b"WWWWWWWWWWW")]
    ins = oracle_inputs(orc, [h], seqs)
    hits, doms = sxa.search_host([h], seqs, *ins, cutoff="bits", seq_t=0.0, dom_t=0.0)
    assert len(hits) == 1 and hits[0]["target"] == 0
    assert doms[hits[0]["dom_first"]]["reported"] == 1


def test_cut_tc_requires_cutoffs(orc, sxa):
    h = _tiny(orc)
    seqs = [orc.encode(b"MKA")]
    ins = oracle_inputs(orc, [h], seqs)
    h.tc = None
    with pytest.raises(sxa.SxError):
        sxa.search_host([h], seqs, *ins, cutoff="tc")
    h.tc = (1.0, 1.0)
    sxa.search_host([h], seqs, *ins, cutoff="tc")


def test_no_pairs_no_hits(orc, sxa):
    h = _tiny(orc)
    hits, doms = sxa.search_host([h], [orc.encode(b"MKA")], [], [], [], [], [], [])
    assert len(hits) == 0 and len(doms) == 0
