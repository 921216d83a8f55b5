"""GPU end-to-end search (sx_search): device cascade, device Forward / Backward with
posterior decoding and region finding, host domain definition from the regions on and
report, against the HMMER 3.4 fixture as the reference
checks it (search.rs:316-355), plus the filtered form and models without STATS lines."""
import os

import numpy as np
import pytest

from test_search_host import GOLDEN, check_hmmer34, read_faa

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def fixture34(orc):
    hmms = orc.parse_hmm_file(os.path.join(GOLDEN, "hmmer34_pair.hmm"))
    tnames, raw = read_faa(os.path.join(GOLDEN, "hmmer34_targets.faa"))
    return hmms, tnames, [orc.encode(s) for s in raw]


def test_search_matches_hmmer34_fixture(gctx, fixture34):
    hmms, tnames, seqs = fixture34
    gctx.set_profiles(hmms)
    gctx.set_targets(seqs)
    hits, doms = gctx.search(cutoff="evalue", seq_e=10.0, dom_e=10.0, filters=False)
    keys = list(zip(hits["profile"], hits["target"]))
    assert keys == sorted(keys)
    check_hmmer34(hits, doms, [h.name for h in hmms], tnames)


def test_search_from_hmm_file_matches_fixture(gctx, fixture34):
    hmms, tnames, seqs = fixture34
    gctx.set_profiles([])
    assert gctx.load_hmm_file(os.path.join(GOLDEN, "hmmer34_pair.hmm")) == 2
    gctx.set_targets(seqs)
    hits, doms = gctx.search(filters=False)
    check_hmmer34(hits, doms, [h.name for h in hmms], tnames)


def test_filtered_search_keeps_scores_of_the_hits_it_reports(gctx, fixture34):
    """With the filters on, a reported hit has the same scores as without them (the
    filters only decide which pairs reach the domain definition)."""
    hmms, tnames, seqs = fixture34
    gctx.set_profiles(hmms)
    gctx.set_targets(seqs)
    h0, d0 = gctx.search(filters=False)
    st0 = dict(gctx.last_search_stats)
    h1, d1 = gctx.search(filters=True)
    st1 = dict(gctx.last_search_stats)
    assert 0 < len(h1) <= len(h0)
    # phase accounting: every pair reaches the host part without filters, fewer with them
    assert st0["n_pairs"] == len(hmms) * len(seqs) and st0["ms_filters"] == 0.0
    assert len(h1) <= st1["n_pairs"] < st0["n_pairs"] and st1["ms_filters"] > 0.0
    assert st0["n_rows"] == len(hmms) * sum(len(x) + 1 for x in seqs)
    assert all(st1[k] >= 0.0 for k in ("ms_specials", "ms_tail", "ms_report"))
    full = {(int(h["profile"]), int(h["target"])): h for h in h0}
    for h in h1:
        g = full[(int(h["profile"]), int(h["target"]))]
        assert h["score"] == g["score"] and h["bias"] == g["bias"] and h["ndom"] == g["ndom"]
        a = d1[h["dom_first"]:h["dom_first"] + h["ndom"]]
        b = d0[g["dom_first"]:g["dom_first"] + g["ndom"]]
        for k in ("hmm_from", "hmm_to", "ali_from", "ali_to", "env_from", "env_to", "score"):
            assert np.array_equal(a[k], b[k])
    strong = [r.split("\t") for r in open(os.path.join(GOLDEN, "hmmer34_expected.tsv")).read().splitlines()]
    names = [h.name for h in hmms]
    found = {(names[h["profile"]], tnames[h["target"]]) for h in h1}
    # every fixture pair above 100 bits passes the filters (6 distinct pairs)
    want = {(f[1], f[0]) for f in strong if float(f[2]) > 100}
    assert len(want) == 6 and want <= found


def test_search_device_and_host_paths_agree(orc, sxa, gctx, fixture34):
    """The device special rows and the oracle's give the same hits (scores within 0.01 bit,
    coordinates equal)."""
    from test_search_host import oracle_inputs
    hmms, tnames, seqs = fixture34
    gctx.set_profiles(hmms)
    gctx.set_targets(seqs)
    hg, dg = gctx.search(filters=False)
    hh, dh = sxa.search_host(hmms, seqs, *oracle_inputs(orc, hmms, seqs), filters=False)
    assert np.array_equal(hg[["profile", "target", "ndom"]], hh[["profile", "target", "ndom"]])
    assert np.abs(hg["score"] - hh["score"]).max() <= 0.01
    for k in ("hmm_from", "hmm_to", "ali_from", "ali_to", "env_from", "env_to"):
        assert np.array_equal(dg[k], dh[k])


def test_device_regions_match_the_host_scan_of_the_same_rows(sxa, gctx, fixture34):
    """Posterior decoding and region finding on the device (sx_search) against the host scan
    (find_regions, through sx_search_host) of the device's own special rows: the same
    regions, so the same hits, domains and coordinates; scores and nexpected differ at
    most by the device's f32 exp / log rounding.  The search copies back only scores and
    the region list, never the special rows."""
    hmms, tnames, seqs = fixture34
    gctx.set_profiles(hmms)
    gctx.set_targets(seqs)
    hg, dg = gctx.search(filters=False)
    st = dict(gctx.last_search_stats)
    pp = np.repeat(np.arange(len(hmms), dtype=np.uint32), len(seqs))
    tt = np.tile(np.arange(len(seqs), dtype=np.uint32), len(hmms))
    fsp, fwd = gctx.specials(sxa.STAGE_FORWARD, pp, tt)
    bsp, bwd = gctx.specials(sxa.STAGE_BACKWARD, pp, tt)
    hh, dh = sxa.search_host(hmms, seqs, pp, tt, fwd, bwd, fsp, bsp, filters=False)
    assert len(hg) >= 6
    assert np.array_equal(hg[["profile", "target", "ndom", "qlen", "tlen"]], hh[["profile", "target", "ndom", "qlen", "tlen"]])
    assert np.abs(hg["score"] - hh["score"]).max() <= 1e-3
    assert np.allclose(hg["nexpected"], hh["nexpected"], rtol=1e-4, atol=1e-5)
    for k in ("hmm_from", "hmm_to", "ali_from", "ali_to", "env_from", "env_to", "reported"):
        assert np.array_equal(dg[k], dh[k]), k
    assert np.abs(dg["score"] - dh["score"]).max() <= 1e-3
    assert st["n_regions"] >= len(hg)
    assert 0 < st["bytes_to_host"] < 4 * st["n_rows"]


def test_uncalibrated_models_skip_the_filters(orc, gctx):
    h = orc.parse_hmm_file(os.path.join(GOLDEN, "tiny.hmm"))[0]
    h.stats = {}  # no STATS lines: score_target runs no filter for it (pipeline.rs:164)
    gctx.set_profiles([h])
    gctx.set_targets([orc.encode(b"GGGGMKAGGGG"), orc.encode(b"WWWWWWWWWWW")])
    hits, doms = gctx.search(cutoff="bits", seq_t=0.0, dom_t=0.0, filters=True)
    assert len(hits) == 1 and hits[0]["target"] == 0 and doms[0]["reported"] == 1
    # and the plain cascade still refuses them
    with pytest.raises(Exception):
        gctx.run_cascade(1)
