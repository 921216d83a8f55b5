"""The allocation-free pair scoring (sx_score_pairs_noalloc) against sx_score_pairs through
the C ABI: the same results, bit for bit (the same tiles and kernels), for every stage and
both launch forms, on two calls in a row of which the second is smaller, with short and
long models; and the size limit."""
import numpy as np
import pytest

from helpers import emit_homolog, random_hmm, random_seq

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def case(orc):
    ms = [7, 33, 64, 150, 260, 513, 1100]
    hmms = [random_hmm(orc, m, 9100 + i) for i, m in enumerate(ms)]
    rng = np.random.default_rng(91)
    seqs = [random_seq(rng, int(rng.integers(1, 600)), star_frac=0.01, degen_frac=0.02) for _ in range(300)]
    seqs += [emit_homolog(orc, hmms[i % len(hmms)], rng, int(rng.integers(60, 500))) for i in range(40)]
    return hmms, seqs


@pytest.mark.parametrize("stage", [0, 1, 2, 3, 4], ids=["msv", "viterbi", "forward", "backward", "msv16"])
@pytest.mark.parametrize("form", [0, 1], ids=["plain", "batched"])
def test_noalloc_matches_score_pairs_twice(sxa, orc, case, stage, form):
    hmms, seqs = case
    if stage == 4:
        hmms = [h for h in hmms if h.m <= 1024]  # the int16 bound's models
    rng = np.random.default_rng(stage * 7 + form)
    with sxa.Context(0) as cx:
        cx.set_profiles(hmms)
        cx.set_targets(seqs)
        for n in (20000, 3001):  # the second call smaller: stale scratch contents must not leak
            prof = rng.integers(0, len(hmms), n).astype(np.uint32)
            tgt = rng.integers(0, len(seqs), n).astype(np.uint32)
            want, _ = cx.score_pairs(stage, prof, tgt, form)
            got, ms = cx.score_pairs_noalloc(stage, prof, tgt, form)
            assert got.tobytes() == want.tobytes(), (n, np.flatnonzero(got != want)[:5])
            assert ms > 0.0


def test_noalloc_limit_and_empty(sxa, orc, case):
    hmms, seqs = case
    with sxa.Context(0) as cx:
        cx.set_profiles(hmms[:2])
        cx.set_targets(seqs[:5])
        n = sxa.NOALLOC_MAX_PAIRS + 1
        with pytest.raises(sxa.SxError):
            cx.score_pairs_noalloc(0, np.zeros(n, np.uint32), np.zeros(n, np.uint32))
        got, _ = cx.score_pairs_noalloc(0, np.zeros(0, np.uint32), np.zeros(0, np.uint32))
        assert len(got) == 0
        with pytest.raises(sxa.SxError):
            cx.score_pairs_noalloc(0, np.array([5], np.uint32), np.array([0], np.uint32))  # profile out of range
        # the full size works
        n = sxa.NOALLOC_MAX_PAIRS
        prof = (np.arange(n) % 2).astype(np.uint32)
        tgt = (np.arange(n) % 5).astype(np.uint32)
        want, _ = cx.score_pairs(1, prof, tgt)
        got, _ = cx.score_pairs_noalloc(1, prof, tgt)
        assert got.tobytes() == want.tobytes()
