"""GPU parity of the device filter cascade (pipeline.rs:159-189) against the oracle's
per-pair cascade (sx_oracle.c sxo_cascade), plus batched vs plain vs graph replay."""
import numpy as np
import pytest

from helpers import emit_homolog, random_hmm, random_seq

pytestmark = pytest.mark.gpu

F = dict(f1=0.3, f2=0.2, f3=0.1)  # loose thresholds so every stage sees work


@pytest.fixture(scope="module")
def calibrated(orc, sxa):
    from sentinel_x_amd import synth
    ms = [12, 40, 90, 150, 230, 333, 520]
    hmms = [random_hmm(orc, m, 4000 + i) for i, m in enumerate(ms)]
    rng = np.random.default_rng(11)
    seqs = []
    for i in range(160):
        if i % 5 == 0:
            seqs.append(emit_homolog(orc, hmms[i % len(hmms)], rng, int(rng.integers(40, 400))))
        else:
            seqs.append(random_seq(rng, int(rng.integers(1, 500)), star_frac=0.005, degen_frac=0.01))
    profs = [orc.configure(h) for h in hmms]

    def scorer(stage, ss):
        b = orc.Batch(profs, ss)
        fn = {0: b.msv, 1: b.viterbi, 2: b.forward}[stage]
        return fn(nthreads=8).reshape(len(profs), len(ss))

    synth.calibrate(hmms, scorer, n=60)
    profs = [orc.configure(h) for h in hmms]
    return hmms, profs, seqs


def oracle_cascade(orc, profs, seqs):
    want = {}
    for p, pr in enumerate(profs):
        for t, s in enumerate(seqs):
            stage, out = orc.cascade(pr, s, **F)
            want[(p, t)] = (stage, out)
    return want


def test_batched_cascade_matches_oracle(orc, gctx, calibrated):
    hmms, profs, seqs = calibrated
    gctx.set_profiles(hmms)
    gctx.set_targets(seqs)
    st, rec = gctx.run_cascade(1, use_graph=False, **F)
    want = oracle_cascade(orc, profs, seqs)
    survivors = {k for k, v in want.items() if v[0] >= 1}
    got = {(int(r["profile"]), int(r["target"])) for r in rec}
    assert got == survivors
    assert st.n_msv_pass == len(survivors)
    assert sum(v[0] >= 2 for v in want.values()) > 5 and sum(v[0] == 4 for v in want.values()) > 2
    same_bias = agree = 0
    for r in rec:
        stage, out = want[(int(r["profile"]), int(r["target"]))]
        assert np.float32(r["msv"]) == out[0]
        if np.float32(r["bias"]) == out[1]:
            same_bias += 1
            assert r["stage"] == stage
            # a record rejected by the int16 Viterbi bound (stage 2) has no f32 score
            if stage >= 3:
                assert not np.isnan(r["vit"])
            if stage >= 2 and not np.isnan(r["vit"]):
                assert np.float32(r["vit"]) == out[2]
        else:
            assert abs(float(r["bias"]) - float(out[1])) <= 1e-4 * max(1.0, abs(float(out[1])))
        agree += int(r["stage"] == stage)
        if stage >= 3 and r["stage"] >= 3:
            f64 = orc.forward64(profs[int(r["profile"])], seqs[int(r["target"])])
            assert abs(float(r["fwd64"]) - f64) <= 1e-3
    assert agree == len(rec)
    assert same_bias >= 0.99 * len(rec)


def test_plain_and_graph_forms_agree_with_batched(gctx, calibrated):
    hmms, profs, seqs = calibrated
    gctx.set_profiles(hmms)
    gctx.set_targets(seqs)
    _, eager = gctx.run_cascade(1, use_graph=False, **F)
    _, plain = gctx.run_cascade(0, **F)
    key = lambda a: np.lexsort((a["target"], a["profile"]))
    eager, plain = eager[key(eager)], plain[key(plain)]
    assert np.array_equal(eager[["profile", "target", "stage"]], plain[["profile", "target", "stage"]])
    for f in ("msv", "bias", "vit"):
        assert np.array_equal(eager[f], plain[f], equal_nan=True)
    assert np.allclose(eager["fwd64"], plain["fwd64"], atol=1e-3, equal_nan=True)
    # graph replays (several, with eager work in between) reproduce the eager run
    for _ in range(3):
        _, g = gctx.run_cascade(1, use_graph=True, **F)
        g = g[key(g)]
        assert np.array_equal(g[["profile", "target", "stage"]], eager[["profile", "target", "stage"]])
        assert np.array_equal(g["vit"], eager["vit"], equal_nan=True)
        gctx.score_all(0, 1)


def test_records_fetch_matches_truncated_run(gctx, calibrated):
    hmms, _, seqs = calibrated
    gctx.set_profiles(hmms)
    gctx.set_targets(seqs)
    same = lambda a, b: a.tobytes() == b.tobytes()  # records hold NaN for stages not reached
    st, full = gctx.run_cascade(1, use_graph=False, **F)
    assert len(full) == st.n_msv_pass > 10
    assert same(gctx.cascade_records(), full)
    assert same(gctx.cascade_records(7), full[:7])
    st2, cut = gctx.run_cascade(1, use_graph=False, cap=5, **F)
    assert st2.n_msv_pass == st.n_msv_pass and len(cut) == 5
    assert same(cut, full[:5])
    _, plain = gctx.run_cascade(0, **F)
    assert same(gctx.cascade_records(), plain)


@pytest.mark.parametrize("bound,vbound", [(False, False), (True, False), (True, True)],
                         ids=["f32", "msv-bound", "msv-and-viterbi-bounds"])
def test_launch_timing_covers_each_stage(gctx, calibrated, bound, vbound):
    hmms, _, seqs = calibrated
    gctx.set_profiles(hmms)
    gctx.set_targets(seqs)
    st, rec = gctx.run_cascade(1, use_graph=False, msv_bound=bound, vit_bound=vbound, **F)
    ls = gctx.cascade_launches()
    assert np.all(ls["ms"] >= 0)
    assert int(ls["cells"][ls["stage"] == 2].sum()) == st.cells_fwd
    if vbound:
        # the bound scores every pair entering Viterbi, the f32 Viterbi its candidates
        assert 5 in set(ls["stage"])
        assert int(ls["cells"][ls["stage"] == 5].sum()) == st.cells_vit
        scored = rec[(rec["stage"] >= 2) & ~np.isnan(rec["vit"])]
        assert len(scored) == st.n_vit_cand >= st.n_vit_pass
        assert st.ms_vit_bound <= st.ms_vit
    else:
        assert int(ls["cells"][ls["stage"] == 1].sum()) == st.cells_vit
    if not bound:
        assert len(ls) >= 3 and set(ls["stage"]) == {0, 1, 2}
        assert int(ls["cells"][ls["stage"] == 0].sum()) == st.cells_msv
        # the MSV launches fill the MSV stage's event window, up to each launch's queue-head
        # reset and launch gap (about 10 us each)
        n0 = int((ls["stage"] == 0).sum())
        assert abs(float(ls["ms"][ls["stage"] == 0].sum()) - st.ms_msv) <= 0.05 + 0.015 * n0 + 0.01 * st.ms_msv
    else:
        # the bound's launches over every pair, then the f32 MSV of its candidates
        assert set(ls["stage"]) - {5} == {4, 0, 1, 2}
        assert int(ls["cells"][ls["stage"] == 4].sum()) == st.cells_msv
        assert int(ls["cells"][ls["stage"] == 0].sum()) == st.cells_msv_exact > 0
        assert st.n_msv_pass <= st.n_msv_cand < st.n_pairs
        n4 = int((ls["stage"] == 4).sum())
        assert abs(float(ls["ms"][ls["stage"] == 4].sum()) - st.ms_msv_bound) <= 0.05 + 0.015 * n4 + 0.01 * st.ms_msv_bound
        assert st.ms_msv_bound <= st.ms_msv
    gctx.run_cascade(1, use_graph=True, msv_bound=bound, vit_bound=vbound, **F)
    assert len(gctx.cascade_launches()) == 0


def test_int16_bound_leaves_every_record_unchanged(gctx, calibrated):
    """The bound only removes pairs the f32 MSV test rejects: with and without it, every
    form gives the same records, and the plain form counts the same candidates."""
    hmms, _, seqs = calibrated
    gctx.set_profiles(hmms)
    gctx.set_targets(seqs)
    key = lambda a: np.lexsort((a["target"], a["profile"]))
    same = lambda a, b: a[key(a)].tobytes() == b[key(b)].tobytes()
    st0, ref = gctx.run_cascade(1, use_graph=False, msv_bound=False, vit_bound=False, **F)
    st1, got = gctx.run_cascade(1, use_graph=False, msv_bound=True, vit_bound=False, **F)
    assert same(ref, got) and st1.n_msv_pass == st0.n_msv_pass
    assert st1.n_msv_cand >= st1.n_msv_pass
    for _ in range(2):
        _, g = gctx.run_cascade(1, use_graph=True, msv_bound=True, vit_bound=False, **F)
        assert same(ref, g)
    stp, plain = gctx.run_cascade(0, msv_bound=True, vit_bound=False, **F)
    _, plain0 = gctx.run_cascade(0, msv_bound=False, vit_bound=False, **F)
    assert same(plain, plain0)
    assert stp.n_msv_cand == st1.n_msv_cand


def test_int16_viterbi_bound_changes_only_unscored_records(gctx, calibrated):
    """With the Viterbi bound every decision and score stays the same; the records the
    bound rejects (stage 2) carry no f32 Viterbi score.  Batched, graph and plain forms
    reject the same pairs."""
    hmms, _, seqs = calibrated
    gctx.set_profiles(hmms)
    gctx.set_targets(seqs)
    key = lambda a: np.lexsort((a["target"], a["profile"]))
    st0, ref = gctx.run_cascade(1, use_graph=False, vit_bound=False, **F)
    st1, got = gctx.run_cascade(1, use_graph=False, vit_bound=True, **F)
    assert st1.n_vit_cand > 0 and st0.n_vit_cand == 0
    ref, got = ref[key(ref)], got[key(got)]
    assert np.array_equal(ref[["profile", "target", "stage"]], got[["profile", "target", "stage"]])
    for f in ("msv", "bias", "fwd64"):
        assert np.array_equal(ref[f], got[f], equal_nan=True)
    nan = np.isnan(got["vit"])
    assert np.all(got["stage"][nan & (ref["stage"] >= 2)] == 2)
    assert np.array_equal(got["vit"][~nan], ref["vit"][~nan])
    assert np.all(np.isnan(ref["vit"][nan]) | (ref["stage"][nan] == 2))
    assert st1.n_vit_pass == st0.n_vit_pass and st1.n_vit_cand >= st1.n_vit_pass
    assert st1.n_vit_cand < st1.n_bias_pass  # the bound rejected some pairs
    for _ in range(2):
        _, g = gctx.run_cascade(1, use_graph=True, vit_bound=True, **F)
        assert g[key(g)].tobytes() == got.tobytes()
    stp, plain = gctx.run_cascade(0, vit_bound=True, **F)
    plain = plain[key(plain)]
    # the plain form's Forward runs other lane shapes (another f32 summation order)
    assert np.array_equal(plain[["profile", "target", "stage"]], got[["profile", "target", "stage"]])
    for f in ("msv", "bias", "vit"):
        assert np.array_equal(plain[f], got[f], equal_nan=True)
    assert np.allclose(plain["fwd64"], got["fwd64"], atol=1e-3, equal_nan=True)
    assert stp.n_vit_cand == st1.n_vit_cand


def test_cascade_requires_calibrated_profiles(orc, gctx, sxa):
    h = random_hmm(orc, 30, 1, stats=False)
    gctx.set_profiles([h])
    gctx.set_targets([np.array([1, 2, 3], np.uint8)])
    with pytest.raises(sxa.SxError):
        gctx.run_cascade(1)
