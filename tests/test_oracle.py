"""Pins the C oracle (oracle/sx_oracle.c) before anything is checked against it.

Sources of truth, per function:
  * the reference's own unit tests, restated (alphabet.rs:104-116, parse.rs:332-370,
    model.rs:364-393, dp.rs:916-955, pipeline.rs:599-604);
  * generic Viterbi / Forward / Backward over an explicit state graph
    (oracle/graph.py) and brute-force path enumeration on the tiniest cases;
  * an independent log-space two-state HMM for the bias filter.
"""
import math
import os

import numpy as np
import pytest

import graph
from helpers import random_hmm, random_seq

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
TINY_TEXT = open(os.path.join(GOLDEN, "tiny.hmm")).read()


# ---------------- alphabet.rs:104-116 ----------------

def test_encodes_canonical_and_degenerate(orc):
    assert list(orc.encode(b"ACDEFGHIKLMNPQRSTVWY")) == list(range(20))
    assert list(orc.encode(b"uo")) == [1, 8]
    assert list(orc.encode(b"BJZX*")) == [20, 21, 22, 23, 24]
    assert list(orc.encode(b"m k\n")) == [10, 8]
    assert list(orc.encode(b"-._~!1")) == [24, 24, 24, 24, 24, 24]
    assert list(orc.encode(b"zq")) == [22, 13]


def test_background_sums_to_one(orc):
    assert abs(float(orc.background().sum()) - 1.0) < 1e-4


# ---------------- parse.rs:332-370 ----------------

def test_parses_header_cutoffs_and_stats(orc):
    h = orc.parse_hmm_text(TINY_TEXT)
    assert len(h) == 1
    h = h[0]
    assert h.name == "tiny" and h.acc == "PF99999.1" and h.m == 3
    assert h.ga == (20.0, 18.5) and h.nc == (10.0, 9.0) and h.tc is None
    assert h.calibrated
    assert abs(h.mat[1][10] - np.float32(math.exp(-0.1))) < 1e-6
    assert h.compo is not None
    assert h.t[3][1] == 0.0


def test_cutoff_pairs(orc):
    assert orc.cutoff_pair("25.00 25.00;") == (25.0, 25.0)
    a, b = orc.cutoff_pair("  792.1 640.3 ;")
    assert (np.float32(a), np.float32(b)) == (np.float32(792.1), np.float32(640.3))
    assert orc.cutoff_pair("30.0 30.0") == (30.0, 30.0)


def test_rejects_nucleotide_models_and_garbage(orc):
    with pytest.raises(ValueError):
        orc.parse_hmm_text(TINY_TEXT.replace("ALPH  amino", "ALPH  DNA"))
    with pytest.raises(ValueError):
        orc.parse_hmm_text(TINY_TEXT.replace("0.02 4.5 4.5 0.6 0.8 0.6 0.8", "0.02 x 4.5 0.6 0.8 0.6 0.8", 1))


def test_parses_two_concatenated_records(orc):
    two = TINY_TEXT + TINY_TEXT.replace("NAME  tiny", "NAME  tiny2")
    h = orc.parse_hmm_text(two)
    assert len(h) == 2 and h[1].name == "tiny2"


# ---------------- model.rs:364-393 ----------------

def flat_hmm(orc, m):
    mat = np.zeros((m + 1, 20), np.float32)
    mat[1:] = orc.background()
    t = np.tile(np.array([0.9, 0.05, 0.05, 0.5, 0.5, 0.5, 0.5], np.float32), (m + 1, 1))
    return orc.Hmm(name="flat", m=m, mat=mat, t=t)


def test_local_entry_sums_to_one(orc):
    p = orc.configure(flat_hmm(orc, 10))
    s = sum(math.exp(p.tsc[7][k]) * (10 - k + 1) for k in range(1, 11))
    assert abs(s - 1.0) < 1e-4


def test_background_match_scores_zero_and_star_neg_inf(orc):
    p = orc.configure(flat_hmm(orc, 5))
    assert np.all(np.abs(p.msc[:24, 1:]) < 1e-5)
    assert np.isneginf(p.msc[24, 1])
    assert np.all(np.isneginf(p.msc[:, 0]))


def test_null_model_matches_hmmer_formula(orc):
    p1 = 300.0 / 301.0
    want = 300.0 * math.log(p1) + math.log(1.0 - p1)
    assert abs(orc.null_score(300) - want) < 1e-3


def test_configure_values_from_first_principles(orc, tiny):
    """msc = ln(p/bg); degenerate = bg-weighted mean; transitions = ln p for nodes 1..M-1;
    D_1 impossible; msv entry = ln(2/(M(M+1)))."""
    h = tiny.hmm
    bg = orc.background().astype(np.float64)
    for k in range(1, h.m + 1):
        for x in range(20):
            assert abs(tiny.msc[x][k] - math.log(h.mat[k][x] / bg[x])) < 1e-5
        xsc = sum(bg[y] * math.log(h.mat[k][y] / bg[y]) for y in range(20)) / bg.sum()
        assert abs(tiny.msc[23][k] - xsc) < 1e-5
        bsc = (bg[2] * math.log(h.mat[k][2] / bg[2]) + bg[11] * math.log(h.mat[k][11] / bg[11])) / (bg[2] + bg[11])
        assert abs(tiny.msc[20][k] - bsc) < 1e-5
    for j in range(7):
        for k in range(1, h.m):
            want = math.log(h.t[k][j]) if h.t[k][j] > 0 else -math.inf
            if j in (5, 6) and k == 1:
                want = -math.inf
            got = tiny.tsc[j][k]
            assert (got == want) if math.isinf(want) else abs(got - want) < 1e-6
        assert np.isneginf(tiny.tsc[j][0]) and np.isneginf(tiny.tsc[j][h.m])
    assert abs(tiny.msv_tbm - math.log(2.0 / (3 * 4))) < 1e-6
    assert np.array_equal(tiny.mo, np.exp(tiny.msc.astype(np.float64)).astype(np.float32)) or \
        np.allclose(tiny.mo, np.exp(tiny.msc.astype(np.float64)), rtol=1e-6)


# ---------------- dp.rs:916-955 ----------------

def test_forward_equals_backward(orc, tiny):
    seq = orc.encode(b"GGMKAGGGMKAGG")
    f = np.float32(orc.forward(tiny, seq))
    b, _ = orc.backward(tiny, seq)
    assert abs(f - b) < 1e-3


def test_viterbi_le_forward(orc, tiny):
    seq = orc.encode(b"GGMKAGGGGGGMKAGG")
    assert orc.viterbi(tiny, seq) <= np.float32(orc.forward(tiny, seq)) + 1e-4


def test_logsum_basics(orc):
    assert abs(orc.logsum(0.0, 0.0) - math.log(2.0)) < 1e-6
    assert orc.logsum(-math.inf, -2.0) == np.float32(-2.0)
    assert orc.logsum(-math.inf, -math.inf) == -math.inf


# ---------------- pipeline.rs:599-604 ----------------

def test_gumbel_and_exp_tails(orc):
    assert abs(orc.gumbel_surv(0.0, 0.0, 1.0) - (1.0 - math.exp(-1.0))) < 1e-12
    assert orc.exp_logsurv(-5.0, 0.0, 0.7) == 0.0
    assert abs(orc.exp_logsurv(10.0, 0.0, 0.7) + 7.0) < 1e-12
    # small-tail branch: 1 - exp(-e^{-y}) ~= e^{-y}
    y = 20.0
    assert abs(orc.gumbel_surv(y, 0.0, 1.0) - math.exp(-y)) < 1e-15


# ---------------- explicit-graph pins ----------------

CASES = [(m, L, seed) for m, L, seed in [(1, 1, 1), (1, 4, 2), (2, 3, 3), (3, 5, 4), (4, 7, 5), (5, 9, 6), (6, 4, 7)]]


def _case(orc, m, L, seed):
    p = orc.configure(random_hmm(orc, m, seed))
    rng = np.random.default_rng(seed + 100)
    d = random_seq(rng, L, star_frac=0.15 if L > 3 else 0.0, degen_frac=0.15)
    return p, d


@pytest.mark.parametrize("m,L,seed", CASES)
def test_msv_matches_generic_max_product(orc, m, L, seed):
    p, d = _case(orc, m, L, seed)
    g = graph.msv_graph(m, p.msc.astype(np.float64), float(p.msv_tbm), L)
    want, _ = g.run(list(d), "S", "T", max)
    got = float(orc.msv(p, d))
    assert (want == got) if math.isinf(want) else abs(got - want) < 1e-4 + 1e-5 * abs(want)


@pytest.mark.parametrize("m,L,seed", CASES)
def test_viterbi_matches_generic_max_product(orc, m, L, seed):
    p, d = _case(orc, m, L, seed)
    g = graph.plan7_graph(m, p.msc.astype(np.float64), p.tsc.astype(np.float64), orc.xsc(L))
    want, _ = g.run(list(d), "S", "T", max)
    got = float(orc.viterbi(p, d))
    assert (want == got) if math.isinf(want) else abs(got - want) < 1e-4 + 1e-5 * abs(want)


def _prob_graph(orc, p, L, multihit=True):
    with np.errstate(divide="ignore"):
        lmo = np.log(p.mo.astype(np.float64))
        ltp = np.log(p.tp.astype(np.float64))
        lx = np.log(orc.xsc(L, multihit)[8:].astype(np.float64))
    return graph.plan7_graph(p.m, lmo, ltp, lx)


SPECIAL_NAMES = ["N", "B", "E", "J", "C"]


@pytest.mark.parametrize("m,L,seed", CASES)
def test_forward_and_specials_match_generic_logsumexp(orc, m, L, seed):
    p, d = _case(orc, m, L, seed)
    g = _prob_graph(orc, p, L)
    want, tab = g.run(list(d), "S", "T", graph._lse)
    f, sp = orc.forward(p, d, specials=True)
    assert abs(f - want) < 1e-5 + 1e-6 * abs(want)
    assert abs(orc.forward64(p, d) - want) < 1e-10 * max(1.0, abs(want))
    for r, name in enumerate(SPECIAL_NAMES):
        for i in range(L + 1):
            w = tab[i][g.index[name]]
            v = float(sp[r][i])
            assert (v == w) if math.isinf(w) else abs(v - w) < 1e-4 + 1e-5 * abs(w), (name, i, v, w)


@pytest.mark.parametrize("m,L,seed", CASES)
def test_backward_and_specials_match_generic_backward(orc, m, L, seed):
    p, d = _case(orc, m, L, seed)
    g = _prob_graph(orc, p, L)
    tab = g.backward(list(d), "T")
    total = tab[0][g.index["S"]]
    b, sp = orc.backward(p, d)
    fwd, _ = g.run(list(d), "S", "T", graph._lse)
    assert abs(total - fwd) < 1e-9 * max(1.0, abs(fwd))
    assert abs(b - total) < 1e-4 + 1e-5 * abs(total)
    assert abs(orc.backward64(p, d) - total) < 1e-6 * max(1.0, abs(total))
    for r, name in enumerate(SPECIAL_NAMES):
        rows = range(1, L + 1) if name not in ("N", "B") else range(0, L + 1)
        for i in rows:
            w = tab[i][g.index[name]]
            v = float(sp[r][i])
            assert (v == w) if math.isinf(w) else abs(v - w) < 1e-4 + 1e-5 * abs(w), (name, i, v, w)


def test_generic_algorithm_equals_path_enumeration(orc):
    p, d = _case(orc, 2, 3, 11)
    g = graph.plan7_graph(2, p.msc.astype(np.float64), p.tsc.astype(np.float64), orc.xsc(3))
    a, _ = g.run(list(d), "S", "T", max)
    b = g.enumerate_paths(list(d), "S", "T", max)
    assert abs(a - b) < 1e-12
    a, _ = g.run(list(d), "S", "T", graph._lse)
    b = g.enumerate_paths(list(d), "S", "T", graph._lse)
    assert abs(a - b) < 1e-12
    gm = graph.msv_graph(2, p.msc.astype(np.float64), float(p.msv_tbm), 3)
    assert abs(gm.run(list(d), "S", "T", max)[0] - gm.enumerate_paths(list(d), "S", "T", max)) < 1e-12


def test_graph_check_catches_a_planted_oracle_mistake(orc):
    """Plant a transposed transition table in the oracle's input (CPU only): the
    explicit-graph check must then disagree, i.e. the pin has teeth."""
    p, d = _case(orc, 5, 9, 6)
    g = graph.plan7_graph(5, p.msc.astype(np.float64), p.tsc.astype(np.float64), orc.xsc(9))
    want, _ = g.run(list(d), "S", "T", max)
    swap = [1, 0, 2, 3, 4, 5, 6, 7]  # MM <-> MI
    bad = orc.Profile(p.name, p.m, p.msc, p.tsc[swap].copy(), p.mo, p.tp[swap].copy(), p.msv_tbm, p.compo, p.hmm)
    assert abs(float(orc.viterbi(bad, d)) - want) > 1e-3
    fwant, _ = _prob_graph(orc, p, 9).run(list(d), "S", "T", graph._lse)
    assert abs(orc.forward(bad, d) - fwant) > 1e-3
    assert abs(orc.forward(p, d) - fwant) < 1e-5


def test_tiny_motif_msv_and_viterbi_against_graph(orc, tiny):
    d = orc.encode(b"GGMKAGGGGGGMKAGG")
    L = len(d)
    want, _ = graph.plan7_graph(3, tiny.msc.astype(np.float64), tiny.tsc.astype(np.float64), orc.xsc(L)).run(list(d), "S", "T", max)
    assert abs(float(orc.viterbi(tiny, d)) - want) < 1e-4
    want, _ = graph.msv_graph(3, tiny.msc.astype(np.float64), float(tiny.msv_tbm), L).run(list(d), "S", "T", max)
    assert abs(float(orc.msv(tiny, d)) - want) < 1e-4


# ---------------- bias filter, cascade ----------------

@pytest.mark.parametrize("seed", [1, 2, 3])
def test_bias_filter_matches_log_space_two_state_hmm(orc, seed):
    p = orc.configure(random_hmm(orc, 40, seed))
    rng = np.random.default_rng(seed)
    d = random_seq(rng, 120, star_frac=0.05, degen_frac=0.1)
    want = graph.two_state_bias(p.compo.astype(np.float64), orc.background().astype(np.float64), p.m, list(d))
    assert abs(float(orc.bias_filter(p, d)) - want) < 1e-4


def test_bias_filter_length_one_by_hand(orc):
    p = orc.configure(random_hmm(orc, 16, 5))
    x = 3
    d = np.array([x], np.uint8)
    bg = orc.background().astype(np.float64)
    want = math.log(0.999 + 0.001 * p.compo[x] / bg[x]) + math.log(0.5) + math.log(0.5)
    assert abs(float(orc.bias_filter(p, d)) - want) < 1e-6


def test_cascade_matches_stagewise_decisions(orc):
    p = orc.configure(random_hmm(orc, 30, 9))
    rng = np.random.default_rng(9)
    LN2 = np.float32(math.log(2.0))
    for _ in range(20):
        d = random_seq(rng, int(rng.integers(20, 200)))
        stage, out = orc.cascade(p, d, f1=0.5, f2=0.5, f3=0.5)
        usc = orc.msv(p, d)
        assert out[0] == usc
        null = np.float32(orc.null_score(len(d)))
        s = p.stats6
        if orc.gumbel_surv(float((usc - null) / LN2), s[0], s[1]) > 0.5:
            assert stage == 0
            continue
        assert stage >= 1 and out[1] == orc.bias_filter(p, d)


def test_empty_inputs_are_neg_inf(orc, tiny):
    e = np.zeros(0, np.uint8)
    assert orc.msv(tiny, e) == -np.inf
    assert orc.viterbi(tiny, e) == -np.inf
    assert orc.forward(tiny, e) == -np.inf


def test_star_only_sequence(orc, tiny):
    d = np.full(6, 24, np.uint8)
    assert orc.msv(tiny, d) == -np.inf or np.isfinite(orc.msv(tiny, d))
    want, _ = graph.msv_graph(3, tiny.msc.astype(np.float64), float(tiny.msv_tbm), 6).run(list(d), "S", "T", max)
    assert want == -math.inf and orc.msv(tiny, d) == -np.inf


# ---------------------------------------------------------------------------
# Quantised MSV bound (the int16 prefilter; no reference counterpart).  It is pinned
# against the f32 MSV, itself pinned above: it must never lie below it, and it may lie
# above it by no more than its roundings allow.
# ---------------------------------------------------------------------------

Q_EMPTY = -2**31


def _q_slack_nats(m, L):
    """Largest gap the bound's roundings allow: each cell of the best segment rounds its
    emission up by at most 1 + 1/32 units of 1/1024 nat, each segment entry rounds B up
    by one unit, each row's N/J arithmetic rounds up by at most 17 fine units
    (2^-22 nat), and a segment may start from the frame floor of the row before, at most
    16 rows of B's fall (|tloop| a row, rounded up) above B."""
    lf = np.float32(L)
    tloop = float(np.float32(math.log(float(lf / (lf + np.float32(3.0))))))
    frame = 16 * math.ceil(abs(tloop) * 1024.0 + 1.0)
    return (min(m, L) * 1.04 + 4.0 + frame) / 1024.0 + L * 20.0 / 4194304.0


def _q_floor_nats(p, L):
    """Where a bound whose rows all peak below B ends: every cell sits at the frame floor
    (B plus at most the frame's slack), so xJ <= B + tEJ and the result is
    B + tEJ + tmove with B <= tmove + tBM (xJ never lifts B: B + tEJ + tmove < B)."""
    lf = np.float32(L)
    tmove = float(np.float32(math.log(float(np.float32(3.0) / (lf + np.float32(3.0))))))
    return 2 * tmove + float(p.msv_tbm) - math.log(2.0)


def q_saturated_floor(orc, L):
    """The least result of a bound one of whose cells reached 32 nats (see sx_oracle.h)."""
    lf = np.float32(L)
    tloop = np.float32(math.log(float(lf / (lf + np.float32(3.0)))))
    tmove = np.float32(math.log(float(np.float32(3.0) / (lf + np.float32(3.0)))))
    c = orc.lib().sxo_q_const
    return 32767 * 4096 + c(np.float32(-0.6931472)) + L * min(c(tloop), 0) + c(tmove)


def test_q_quantisation_rounds_up(orc):
    L = orc.lib()
    assert L.sxo_q_emit(np.float32(-np.inf)) == -32768
    for v in (-40.0, -33.0, -5.3, -1e-3, 0.0, 1e-7, 0.25, 3.9, 40.0):
        q = L.sxo_q_emit(np.float32(v))
        assert q >= max(-32768, min(32767, math.ceil(1024.0 * float(np.float32(v)))))
        assert q <= max(-32768, min(32767, math.ceil(1024.0 * float(np.float32(v))) + 1))
    for v in (-9.4, -0.6931472, -1e-4, 0.0, 2.0):
        c = L.sxo_q_const(np.float32(v))
        assert 4194304.0 * float(np.float32(v)) + 16 <= c < 4194304.0 * float(np.float32(v)) + 17


@pytest.mark.parametrize("m", [1, 2, 7, 40, 150, 400, 1024])
def test_msv16_bounds_f32_msv_from_above(orc, m):
    from helpers import emit_homolog
    rng = np.random.default_rng(7000 + m)
    h = random_hmm(orc, m, 7100 + m)
    p = orc.configure(h)
    F = orc.Q_UNITS_PER_NAT
    n = 0
    for L in (1, 2, 3, 17, 64, 300, 1200, 4000):
        for kind in range(4):
            if kind == 2:
                d = emit_homolog(orc, h, rng, L)
            elif kind == 3:
                # two hits apart: xJ lifts B mid-target, so the frame floor moves up
                a = emit_homolog(orc, h, rng, max(1, L // 3))
                d = np.concatenate([random_seq(rng, L // 3, 0.0, 0.0), a, random_seq(rng, L // 6, 0.0, 0.0),
                                    emit_homolog(orc, h, rng, max(1, L // 4))]).astype(np.uint8)
            else:
                d = random_seq(rng, L, star_frac=0.02 * kind, degen_frac=0.03 * kind)
            u, q = orc.msv(p, d), orc.msv16(p, d)
            assert q != Q_EMPTY
            floor_sat = q_saturated_floor(orc, len(d))
            # a bound whose cells saturated ends at or above floor_sat; below it none did
            assert q >= min(math.floor(F * float(u)), floor_sat), (L, kind, q, float(u))
            if q >= floor_sat:
                continue
            assert q >= math.floor(F * float(u)), (L, kind, q, float(u))
            # tight, unless the rows peak below B, where the floor holds every cell
            tight = q / F - float(u) <= _q_slack_nats(m, len(d))
            assert tight or q / F <= _q_floor_nats(p, len(d)) + _q_slack_nats(m, len(d)), \
                (L, kind, q / F - float(u), q / F)
            n += tight
    assert n >= 14


def test_msv16_saturates_above_its_floor_and_is_empty_for_l0(orc):
    m = 200
    mat = np.full((m + 1, 20), 0.01 / 19, np.float32)
    mat[:, 18] = 0.99  # W at every position
    t = np.tile(np.array([0.9, 0.05, 0.05, 0.7, 0.3, 0.7, 0.3], np.float32), (m + 1, 1))
    t[m] = [1.0, 0.0, 0.0, 1.0, 0.0, 1.0, 0.0]
    p = orc.configure(orc.Hmm(name="w", m=m, mat=mat, t=t, compo=mat[1:].mean(axis=0)))
    for L in (30, 200, 3000):
        d = np.full(L, 18, np.uint8)
        assert float(orc.msv(p, d)) > 32.0
        assert orc.msv16(p, d) >= q_saturated_floor(orc, L)
    assert orc.msv16(p, np.zeros(0, np.uint8)) == Q_EMPTY
    # a star-only target: the f32 score is -inf; every cell sits at the frame floor, which
    # starts at B = tmove + tBM and falls with xN by at most tloop a row, so the bound is
    # B + tEJ + tmove less at most (L - 1) |tloop|
    L = 30
    q = orc.msv16(p, np.full(L, 24, np.uint8)) / orc.Q_UNITS_PER_NAT
    hi = _q_floor_nats(p, L)
    tloop = math.log(L / (L + 3.0))
    assert hi + (L - 1) * tloop - 0.01 <= q <= hi + 0.01, (q, hi)


def test_msv16_batch_driver_matches_single_pair(orc):
    rng = np.random.default_rng(5)
    hm = [random_hmm(orc, m, 40 + m) for m in (3, 60, 129)]
    profs = [orc.configure(h) for h in hm]
    seqs = [random_seq(rng, L, star_frac=0.01, degen_frac=0.02) for L in (0, 1, 33, 250)]
    got = orc.Batch(profs, seqs).msv16(nthreads=4).reshape(len(profs), len(seqs))
    for i, p in enumerate(profs):
        for j, d in enumerate(seqs):
            assert got[i, j] == orc.msv16(p, d)


# ---------------------------------------------------------------------------
# Quantised Viterbi bound (the int16 Viterbi prefilter; no reference counterpart),
# pinned against the f32 Viterbi the same way as the MSV bound.
# ---------------------------------------------------------------------------

def v_saturated_floor(orc, L):
    """The least result of a Viterbi bound one of whose cells reached 32 nats."""
    xs = orc.xsc(L)
    c = orc.lib().sxo_q_const
    return 32767 * 4096 + c(np.float32(-0.6931472)) + L * min(c(xs[0]), 0) + c(xs[1])


def _v_slack_nats(m, L):
    """Each cell of the best path rounds its emission and transition up (at most 2 * 1.04
    units of 1/1024 nat), each row rounds B up by one unit and N/J by 17 fine units."""
    return (min(m, L) * 2.1 + L * 0.02 + 6.0) / 1024.0 + L * 20.0 / 4194304.0


def test_q_trans_keeps_nonpositive_scores_nonpositive(orc):
    L = orc.lib()
    assert L.sxo_q_trans(np.float32(-np.inf)) == -32768
    for v in (-1e-9, -1e-6, -1e-5, 0.0):
        assert L.sxo_q_trans(np.float32(v)) == 0
    for v in (-3.0, -0.105, 1e-6, 0.5):
        q = L.sxo_q_trans(np.float32(v))
        assert q >= math.ceil(1024.0 * float(np.float32(v)) - 1e-9) and q <= math.ceil(1024.0 * float(np.float32(v))) + 1


@pytest.mark.parametrize("m", [1, 2, 7, 40, 150, 400, 1024])
def test_vit16_bounds_f32_viterbi_from_above(orc, m):
    from helpers import emit_homolog
    rng = np.random.default_rng(9000 + m)
    h = random_hmm(orc, m, 9100 + m)
    p = orc.configure(h)
    F = orc.Q_UNITS_PER_NAT
    n = 0
    for L in (1, 2, 3, 17, 64, 300, 1200):
        for kind in range(3):
            if kind == 2:
                d = emit_homolog(orc, h, rng, L)
            else:
                d = random_seq(rng, L, star_frac=0.02 * kind, degen_frac=0.03 * kind)
            v, q = orc.viterbi(p, d), orc.vit16(p, d)
            floor_sat = v_saturated_floor(orc, len(d))
            assert q >= min(math.floor(F * float(v)), floor_sat), (L, kind, q, float(v))
            if q >= floor_sat:
                continue
            assert q >= math.floor(F * float(v)), (L, kind, q, float(v))
            assert q / F - float(v) <= _v_slack_nats(m, len(d)), (L, kind, q / F - float(v))
            n += 1
    assert n >= 14


def test_vit16_batch_driver_and_edges(orc):
    rng = np.random.default_rng(6)
    hm = [random_hmm(orc, m, 60 + m) for m in (1, 3, 60, 129)]
    profs = [orc.configure(h) for h in hm]
    seqs = [random_seq(rng, L, star_frac=0.05, degen_frac=0.02) for L in (0, 1, 33, 250)]
    got = orc.Batch(profs, seqs).vit16(nthreads=4).reshape(len(profs), len(seqs))
    for i, p in enumerate(profs):
        for j, d in enumerate(seqs):
            assert got[i, j] == orc.vit16(p, d)
    assert np.all(got[:, 0] == -2**31)
