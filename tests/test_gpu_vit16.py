"""GPU parity of the int16 Viterbi bound (SX_STAGE_VIT16, the Viterbi stage's prefilter; no
reference counterpart) with the oracle's sxo_vit16: bit-exact for every kernel shape, in
the all-pairs and pair-list forms.  That the bound never lies below the f32 Viterbi score
is pinned on the CPU (test_oracle.py) and end to end by the cascade tests."""
import numpy as np
import pytest

from helpers import random_hmm
from test_gpu_stages import make_targets
from test_oracle import v_saturated_floor

pytestmark = pytest.mark.gpu

Q_EMPTY = -2**31

# (Q, G, table mode): one-lane groups (0) and aligned lane blocks of 8, 16, 32 lanes (3)
SHAPES = [(16, 1, 0), (32, 1, 0)] + [(q, g, 3) for g in (8, 16, 32) for q in (16, 24, 32)]


@pytest.mark.parametrize("q,g,h", SHAPES, ids=[f"Q{q}G{g}M{h}" for q, g, h in SHAPES])
def test_every_vit16_shape_bit_exact(orc, gctx, q, g, h):
    mmax = min(q * g, 1024)
    rng = np.random.default_rng(q * 13 + g * 5 + h)
    lo = max(1, (q * g) // 2)
    ms = sorted({lo, (lo + mmax) // 2, mmax, int(rng.integers(lo, mmax + 1))})
    hmms = [random_hmm(orc, m, 9800 + q * 131 + g * 7 + h + i) for i, m in enumerate(ms)]
    seqs = make_targets(orc, hmms, 300 if g == 1 else 150, 9800 + q + g + h, lmax=220, homolog_frac=0.2)
    profs = [orc.configure(x) for x in hmms]
    gctx.set_profiles(hmms)
    gctx.set_targets(seqs)
    gctx.set_shapes(5, [(mmax, q, g, h)])
    got, _ = gctx.score_all(5, 1)
    pairs = (np.arange(60) % len(hmms), (np.arange(60) * 7) % len(seqs))
    gp, _ = gctx.score_pairs(5, *pairs)
    gctx.set_shapes(5, [])
    want = orc.Batch(profs, seqs).vit16(nthreads=8).reshape(len(profs), len(seqs))
    assert np.array_equal(got, want), np.argwhere(got != want)[:5]
    assert np.array_equal(gp, want[pairs])


@pytest.mark.parametrize("form", [0, 1], ids=["plain", "batched"])
def test_vit16_default_shapes_and_saturation(orc, gctx, form):
    ms = [1, 2, 5, 16, 17, 32, 33, 100, 128, 129, 255, 256, 400, 511, 700, 1023, 1024]
    hmms = [random_hmm(orc, m, 9900 + i) for i, m in enumerate(ms)]
    m = 200
    mat = np.full((m + 1, 20), 0.01 / 19, np.float32)
    mat[:, 18] = 0.99
    t = np.tile(np.array([0.9, 0.05, 0.05, 0.7, 0.3, 0.7, 0.3], np.float32), (m + 1, 1))
    t[m] = [1.0, 0.0, 0.0, 1.0, 0.0, 1.0, 0.0]
    hmms.append(orc.Hmm(name="w", m=m, mat=mat, t=t, compo=mat[1:].mean(axis=0)))
    seqs = make_targets(orc, hmms, 90, 9901, lmax=400, homolog_frac=0.2)
    seqs += [np.zeros(0, np.uint8), np.full(180, 18, np.uint8)]
    profs = [orc.configure(x) for x in hmms]
    gctx.set_profiles(hmms)
    gctx.set_targets(seqs)
    gctx.set_shapes(5, [])
    got, _ = gctx.score_all(5, form)
    want = orc.Batch(profs, seqs).vit16(nthreads=8).reshape(len(profs), len(seqs))
    assert np.array_equal(got, want), np.argwhere(got != want)[:5]
    assert np.all(got[:, -2] == Q_EMPTY)
    assert float(orc.viterbi(profs[-1], seqs[-1])) > 32.0
    assert got[-1, -1] >= v_saturated_floor(orc, len(seqs[-1]))
