"""GPU streamed search over a model file (sx_search_hmm_file, search_hmm_file_models of
search.rs:252-271): the file is searched a batch at a time while host threads parse and
configure the next batches.  The hits must not depend on the batch size and must equal
one search over every model of the file loaded at once (sx_load_hmm_file + sx_search);
a packed profile cache gives the same; the HMMER 3.4 fixture reads back through it; hits
refer to models by their index in the file, whose identities come back in file order."""
import os

import numpy as np
import pytest

from helpers import calibrate_with_oracle
from test_search_host import GOLDEN, check_hmmer34, read_faa

pytestmark = pytest.mark.gpu

NTH = min(16, os.cpu_count() or 1)
F = dict(f1=0.02, f2=1e-3, f3=1e-5)


@pytest.fixture(scope="module")
def marker_file(orc, tmp_path_factory):
    """40 marker-set profiles (calibrated by the oracle) written as an HMMER3 file, with the
    marker-set proteins."""
    from sentinel_x_amd import synth
    wl = synth.make_workload(1)
    nhom = np.bincount(wl.homolog_of[wl.homolog_of >= 0], minlength=len(wl.hmms))
    pick = [int(i) for i in np.argsort(-nhom, kind="stable")[:40]]
    hmms = [wl.hmms[i] for i in sorted(pick)]
    for i, h in enumerate(hmms):
        h.acc = f"PF{i:05d}.9" if i % 2 else None
    calibrate_with_oracle(orc, synth, hmms, NTH)
    d = tmp_path_factory.mktemp("gingest")
    path = str(d / "markers.hmm.gz")
    synth.write_hmm(path, hmms)
    return path, hmms, wl.seqs


def _same(a, b):
    (h1, d1), (h2, d2) = a, b
    assert np.array_equal(h1, h2)
    assert np.array_equal(d1, d2)


@pytest.mark.parametrize("filters", [True, False])
def test_batches_give_the_search_of_the_whole_file(gctx, marker_file, filters):
    path, hmms, seqs = marker_file
    gctx.set_targets(seqs)
    gctx.set_profiles([])
    assert gctx.load_hmm_file(path) == len(hmms)
    kw = dict(cutoff="evalue", seq_e=10.0, dom_e=10.0, filters=filters, **F)
    whole = gctx.search(**kw)
    assert len(whole[0]) >= 10
    for batch in (1, 7, 0, 64):
        h, d, models = gctx.search_hmm_file(path, batch=batch, **kw)
        _same((h, d), whole)
        assert [m[0] for m in models] == [x.name for x in hmms]
        assert [m[1] for m in models] == [x.acc for x in hmms]
        assert [m[2] for m in models] == [x.m for x in hmms]
        st = gctx.last_search_stats
        assert st["n_models"] == len(hmms) and st["ms_ingest"] >= 0.0
    assert gctx.n_profiles == len(hmms)  # the last batch (64 > 40 models: all of them)


def test_profile_cache_gives_the_same_search(sxa, gctx, marker_file, tmp_path):
    path, hmms, seqs = marker_file
    gctx.set_targets(seqs)
    gctx.set_profiles([])
    gctx.load_hmm_file(path)
    cache = str(tmp_path / "markers.sxpc")
    gctx.write_profile_cache(cache)
    conv = str(tmp_path / "conv.sxpc")
    assert sxa.convert_hmm_file(path, conv) == len(hmms)
    assert open(cache, "rb").read() == open(conv, "rb").read()
    kw = dict(cutoff="ga", filters=True, **F)
    ref = gctx.search_hmm_file(path, batch=16, **kw)
    got = gctx.search_hmm_file(cache, batch=16, **kw)
    _same(ref[:2], got[:2])
    assert ref[2] == got[2]
    gctx.set_profiles([])
    assert gctx.load_hmm_file(cache) == len(hmms)
    _same(gctx.search(**kw), ref[:2])


def test_hmmer34_fixture_through_the_stream(orc, sxa, gctx, tmp_path):
    src = os.path.join(GOLDEN, "hmmer34_pair.hmm")
    tnames, raw = read_faa(os.path.join(GOLDEN, "hmmer34_targets.faa"))
    gctx.set_targets([orc.encode(s) for s in raw])
    cache = str(tmp_path / "pair.sxpc")
    sxa.convert_hmm_file(src, cache)
    for path in (src, cache):
        for batch in (1, 0):
            hits, doms, models = gctx.search_hmm_file(path, batch=batch, cutoff="evalue", seq_e=10.0, dom_e=10.0,
                                                      filters=False)
            check_hmmer34(hits, doms, [m[0] for m in models], tnames)


def test_parse_error_reaches_the_caller(sxa, gctx, marker_file, tmp_path):
    path, hmms, seqs = marker_file
    bad = str(tmp_path / "bad.hmm")
    open(bad, "w").write(open(os.path.join(GOLDEN, "tiny.hmm")).read().replace("LENG  3", "LENG  x"))
    with pytest.raises(sxa.SxError, match=r"line 5: bad LENG value"):
        gctx.search_hmm_file(bad, cutoff="evalue")
