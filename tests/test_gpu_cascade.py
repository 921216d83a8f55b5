"""GPU parity of the device filter cascade (pipeline.rs:159-189) against the oracle's
per-pair cascade (sx_oracle.c sxo_cascade), plus batched vs plain vs graph replay."""
import os

import numpy as np
import pytest

from conftest import GOLDEN
from helpers import check_cascade_records, emit_homolog, fwd_boundaries, random_hmm, random_seq

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
    for r in rec:
        stage, out = want[(int(r["profile"]), int(r["target"]))]
        assert np.float32(r["msv"]) == out[0]
        assert np.float32(r["bias"]) == out[1]
        assert r["stage"] == stage
        if stage >= 2:
            assert np.float32(r["vit"]) == out[2]
        if stage >= 3:
            f64 = orc.forward64(profs[int(r["profile"])], seqs[int(r["target"])])
            assert abs(float(r["fwd64"]) - f64) <= 1e-3
    assert st.n_bias_recheck == int(np.sum(rec["flags"] & 1))


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
def test_launch_records_cover_each_stage(gctx, calibrated, bound, vbound):
    """Structure of the per-launch records: the stages launched and the cells each stage's
    launches scored.  (How the launch times fill the stage windows is a timing figure,
    reported by bench.py, not checked here.)"""
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
    else:
        assert int(ls["cells"][ls["stage"] == 1].sum()) == st.cells_vit
    if not bound:
        assert len(ls) >= 3 and set(ls["stage"]) == {0, 1, 2}
        assert int(ls["cells"][ls["stage"] == 0].sum()) == st.cells_msv
    else:
        # the bound's launches over every pair, then the f32 MSV of its candidates
        assert set(ls["stage"]) - {5} == {4, 0, 1, 2}
        assert int(ls["cells"][ls["stage"] == 4].sum()) == st.cells_msv
        assert int(ls["cells"][ls["stage"] == 0].sum()) == st.cells_msv_exact > 0
        assert st.n_msv_pass <= st.n_msv_cand < st.n_pairs
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


def test_every_bias_score_through_the_host_recheck(orc, sxa, gctx, calibrated, monkeypatch):
    """With SX_BIAS_RECHECK_ALL=1 the device hands every bias score to the host, which
    recomputes it in the reference's order: the records (scores, stages, everything but the
    flag) are those of the default run, in every form."""
    hmms, profs, seqs = calibrated
    gctx.set_profiles(hmms)
    gctx.set_targets(seqs)
    key = lambda a: np.lexsort((a["target"], a["profile"]))
    runs = {}
    for form, graph in ((1, False), (1, True), (0, False)):
        st, r = gctx.run_cascade(form, use_graph=graph, **F)
        runs[(form, graph)] = (st, r[key(r)])
    monkeypatch.setenv("SX_BIAS_RECHECK_ALL", "1")
    with sxa.Context(0) as cx:
        cx.set_profiles(hmms)
        cx.set_targets(seqs)
        for (form, graph), (st0, r0) in runs.items():
            st, r = cx.run_cascade(form, use_graph=graph, **F)
            r = r[key(r)]
            assert st.n_bias_recheck == st.n_msv_pass == len(r)
            assert np.all(r["flags"] & sxa.REC_BIAS_HOST)
            a, b = r.copy(), r0.copy()
            a["flags"] = b["flags"] = 0
            assert a.tobytes() == b.tobytes(), (form, graph)
            assert (st.n_bias_pass, st.n_vit_pass, st.n_fwd_pass) == (st0.n_bias_pass, st0.n_vit_pass,
                                                                      st0.n_fwd_pass)


@pytest.mark.parametrize("overflow", [False, True], ids=["several-blocks", "list-overflow"])
def test_compaction_across_blocks_and_list_growth(orc, gctx, overflow):
    """7 profiles x 12,000 short targets (21 compaction blocks of 4,096 pairs).  With
    calibrated STATS the survivors come from many blocks; with an MSV location far below
    every score every pair survives MSV, the bound's candidate and survivor lists (65,536
    entries at first) overflow in a later block and are grown, and the cascade runs again.
    The records equal the oracle's cascade and come out in (profile position, length-sorted
    target) order."""
    from sentinel_x_amd import synth
    ms = [12, 30, 55, 80, 110, 140, 170]
    hmms = [random_hmm(orc, m, 4100 + i) for i, m in enumerate(ms)]
    rng = np.random.default_rng(4100)
    seqs = []
    for i in range(12000):
        L = int(rng.integers(20, 150))
        seqs.append(emit_homolog(orc, hmms[i % 7], rng, L) if i % 9 == 0 else random_seq(rng, L, degen_frac=0.01))
    profs = [orc.configure(h) for h in hmms]

    def scorer(stage, ss):
        b = orc.Batch(profs, ss)
        return {0: b.msv, 1: b.viterbi, 2: b.forward}[stage](nthreads=8).reshape(len(profs), len(ss))

    synth.calibrate(hmms, scorer, n=60)
    if overflow:
        for h in hmms:
            h.stats = dict(h.stats, MSV=(-1.0e4, h.stats["MSV"][1]))
    profs = [orc.configure(h) for h in hmms]
    f = dict(f1=0.02, f2=0.05, f3=0.01)
    stage_o, sc_o = orc.Batch(profs, seqs).cascade(**f, nthreads=8)
    gctx.set_profiles(hmms)
    gctx.set_targets(seqs)
    nt = len(seqs)
    pos = np.argsort(np.argsort(np.array(ms), kind="stable"), kind="stable")  # profile position by M
    lens = np.array([len(x) for x in seqs])
    spos = np.empty(nt, np.int64)
    spos[np.argsort(-lens, kind="stable")] = np.arange(nt)  # targets by decreasing length
    fb = fwd_boundaries(hmms, f["f3"])
    for bound in (True, False):
        for graph in (False, True):
            st, rec = gctx.run_cascade(1, use_graph=graph, msv_bound=bound, **f)
            order = pos[rec["profile"]].astype(np.int64) * nt + spos[rec["target"]]
            assert np.all(np.diff(order) > 0), "records out of (profile position, target) order"
            check_cascade_records(rec, stage_o, sc_o, nt, 1e-3, fb)
            if overflow:
                assert st.n_msv_pass == len(profs) * nt > 65536
    assert int((stage_o >= 2).sum()) > 100 and int((stage_o == 4).sum()) > 10


def test_threshold_adjacent_pairs_decide_as_the_reference(orc, gctx):
    """Pairs placed half an f32 ulp either side of each test's boundary, and exactly on it
    (tests/cascade_pins.py): every device decision is the one the construction fixes.  The
    MSV, bias and Viterbi boundaries are placed with the oracle's scores (the device's are
    bit-identical); the Forward boundary with the device's own Forward score (the contract
    for Forward is the FP64 tolerance, so only the device's score fixes its side)."""
    import cascade_pins as cp
    rng = np.random.default_rng(78)
    base = [random_hmm(orc, m, 4200 + i) for i, m in enumerate((40, 130))]
    seqs = []
    for i in range(16):
        L = int(rng.integers(30, 400))
        seqs.append(cp.biased_seq(rng, orc.configure(base[i % 2]).compo, L) if i % 4 >= 2 else
                    random_seq(rng, L, degen_frac=0.02))
    # device Forward scores for every (base model, target), all filters open, per form:
    # the plain form's Forward runs other lane shapes (another summation order)
    loose = {"MSV": (cp.LOOSE, cp.LAM), "VITERBI": (cp.LOOSE, cp.LAM), "FORWARD": (cp.LOOSE, cp.LAM)}
    for h in base:
        h.stats = dict(loose)
    gctx.set_profiles(base)
    gctx.set_targets(seqs)
    ncases = 0
    for form, graphs in ((1, (False, True)), (0, (False,))):
        _, r0 = gctx.run_cascade(form, use_graph=False, f1=0.5, f2=0.5, f3=0.5)
        dev_fwd = {(int(r["profile"]), int(r["target"])): np.float32(r["fwd"]) for r in r0}
        # one model copy per (base, target, case), grouped by the case's thresholds
        groups = {}
        for b, h in enumerate(base):
            p = orc.configure(h)
            for t, d in enumerate(seqs):
                if t % 2 != b:
                    continue
                sc = (orc.msv(p, d), np.float32(orc.null_score(len(d))), orc.bias_filter(p, d), orc.viterbi(p, d),
                      dev_fwd[(b, t)])
                if cp.ulps_apart(cp.bits(sc[0], sc[1]), cp.bits(sc[0], sc[2])) < 8:
                    continue
                for name, st6, fs, want in cp.cases(*sc):
                    c = orc.Hmm(name=f"{h.name}_{t}_{name}", m=h.m, mat=h.mat, t=h.t, compo=h.compo,
                                stats={"MSV": tuple(st6[0:2]), "VITERBI": tuple(st6[2:4]),
                                       "FORWARD": tuple(st6[4:6])})
                    groups.setdefault(fs, []).append((c, t, want, name))
        for fs, items in groups.items():
            gctx.set_profiles([c for c, _, _, _ in items])
            for graph in graphs:
                _, rec = gctx.run_cascade(form, use_graph=graph, f1=fs[0], f2=fs[1], f3=fs[2])
                got = {(int(r["profile"]), int(r["target"])): int(r["stage"]) for r in rec}
                for i, (_, t, want, name) in enumerate(items):
                    assert got.get((i, t), 0) == want, (name, t, form, graph, fs)
                    ncases += 1
        gctx.set_profiles(base)
    assert ncases >= 3 * 8 * 11


def test_profile_info_gives_model_identity(orc, gctx):
    """ModelInfo (search.rs:245-249) for models loaded from a HMMER3 file, in file order."""
    path = os.path.join(GOLDEN, "hmmer34_pair.hmm")
    gctx.set_profiles([])
    n = gctx.load_hmm_file(path)
    want = orc.parse_hmm_file(path)
    assert n == len(want) == 2
    for i, h in enumerate(want):
        assert gctx.profile_info(i) == (h.name, h.acc, h.desc, h.m)
