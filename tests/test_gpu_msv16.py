"""GPU parity of the int16 MSV bound (SX_STAGE_MSV16, the MSV stage's prefilter; no
reference counterpart) with the oracle's sxo_msv16: bit-exact for every kernel shape, in
the all-pairs and pair-list forms.  That the bound never lies below the f32 MSV is pinned
on the CPU (test_oracle.py) and end to end by the cascade tests (identical records with
and without the bound)."""
import math

import numpy as np
import pytest

from helpers import random_hmm
from test_gpu_stages import make_targets
from test_oracle import q_saturated_floor

pytestmark = pytest.mark.gpu

Q_EMPTY = -2**31
Q_MAX_M = 1024

# (Q positions per lane, G lanes per pair, table mode): one-lane groups (0) and aligned
# lane blocks of 4, 8, 16, 32 lanes (3: 512-thread, 4: 1024-thread workgroups)
SHAPES = [(q, 1, 0) for q in (16, 24, 32, 48, 64, 80, 96, 112, 128, 144, 160, 176, 192, 208, 224, 240, 256,
                              288, 320)] + \
    [(q, 8, h) for q in (16, 24, 32, 40, 48, 56, 64, 72, 80) for h in (3, 4)] + \
    [(88, 8, 3), (96, 8, 3), (128, 8, 3)] + \
    [(q, 4, h) for q in (72, 80, 88, 96, 104, 112, 128) for h in (3, 4)] + \
    [(q, 16, h) for q in (16, 24, 32, 40, 48, 56, 64) for h in (3, 4)] + \
    [(q, 32, h) for q in (16, 24, 32) for h in (3, 4)]


@pytest.mark.parametrize("q,g,h", SHAPES, ids=[f"Q{q}G{g}M{h}" for q, g, h in SHAPES])
def test_every_msv16_shape_bit_exact(orc, gctx, q, g, h):
    """Models from half the shape's capacity to its capacity (the model sits at the right
    end of the lane blocks, after the padding slots), targets for several tiles per
    profile, the last partial."""
    mmax = min(q * g, Q_MAX_M)
    rng = np.random.default_rng(q * 11 + g * 3 + h)
    lo = max(1, (q * g) // 2)
    ms = sorted({lo, (lo + mmax) // 2, mmax, int(rng.integers(lo, mmax + 1))})
    hmms = [random_hmm(orc, m, 8000 + q * 131 + g * 7 + h + i) for i, m in enumerate(ms)]
    seqs = make_targets(orc, hmms, 300 if g == 1 else 150, 8000 + q + g + h, lmax=220, homolog_frac=0.2)
    profs = [orc.configure(x) for x in hmms]
    gctx.set_profiles(hmms)
    gctx.set_targets(seqs)
    gctx.set_shapes(4, [(mmax, q, g, h)])
    got, _ = gctx.score_all(4, 1)
    pairs = (np.arange(60) % len(hmms), (np.arange(60) * 7) % len(seqs))
    gp, _ = gctx.score_pairs(4, *pairs)
    gctx.set_shapes(4, [])
    want = orc.Batch(profs, seqs).msv16(nthreads=8).reshape(len(profs), len(seqs))
    assert np.array_equal(got, want), np.argwhere(got != want)[:5]
    assert np.array_equal(gp, want[pairs])


@pytest.mark.parametrize("form", [0, 1], ids=["plain", "batched"])
def test_msv16_default_shapes_and_saturation(orc, gctx, form):
    """Built-in shapes over model lengths 1..1024, with empty targets (INT32_MIN) and a
    strong hit whose cells saturate the int16 range."""
    ms = [1, 2, 5, 17, 63, 64, 127, 128, 200, 255, 256, 400, 511, 700, 1023, 1024]
    hmms = [random_hmm(orc, m, 8500 + i) for i, m in enumerate(ms)]
    m = 200
    mat = np.full((m + 1, 20), 0.01 / 19, np.float32)
    mat[:, 18] = 0.99
    t = np.tile(np.array([0.9, 0.05, 0.05, 0.7, 0.3, 0.7, 0.3], np.float32), (m + 1, 1))
    t[m] = [1.0, 0.0, 0.0, 1.0, 0.0, 1.0, 0.0]
    hmms.append(orc.Hmm(name="w", m=m, mat=mat, t=t, compo=mat[1:].mean(axis=0)))
    seqs = make_targets(orc, hmms, 90, 8501, lmax=400, homolog_frac=0.2)
    seqs += [np.zeros(0, np.uint8), np.full(180, 18, np.uint8)]
    profs = [orc.configure(x) for x in hmms]
    gctx.set_profiles(hmms)
    gctx.set_targets(seqs)
    gctx.set_shapes(4, [])
    got, _ = gctx.score_all(4, form)
    want = orc.Batch(profs, seqs).msv16(nthreads=8).reshape(len(profs), len(seqs))
    assert np.array_equal(got, want), np.argwhere(got != want)[:5]
    assert np.all(got[:, -2] == Q_EMPTY)
    assert float(orc.msv(profs[-1], seqs[-1])) > 32.0
    assert got[-1, -1] >= q_saturated_floor(orc, len(seqs[-1]))


def test_msv16_rejects_models_beyond_its_shapes(orc, gctx, sxa):
    gctx.set_profiles([random_hmm(orc, 1100, 8600)])
    gctx.set_targets([np.array([1, 2, 3], np.uint8)])
    with pytest.raises(sxa.SxError):
        gctx.score_all(4, 1)
    with pytest.raises(sxa.SxError):
        gctx.set_shapes(4, [(1100, 128, 8, 3)])
    with pytest.raises(sxa.SxError):
        gctx.set_shapes(4, [(300, 32, 8, 3)])  # capacity 256 below its max_m


def test_msv16_long_multi_hit_targets(orc, gctx):
    """Targets of 2,000-40,000 residues, some with hits spread along them (xJ lifts B
    several times, so the frame floor moves up as well as down on its 16-row schedule),
    on the default shapes and on a one-lane, a four-lane and an eight-lane shape: the
    bound is the oracle's, bit for bit, and it never lies below the f32 MSV."""
    from helpers import emit_homolog, random_seq
    rng = np.random.default_rng(8700)
    ms = [60, 150, 300, 450]
    hmms = [random_hmm(orc, m, 8700 + i) for i, m in enumerate(ms)]
    seqs = []
    for L in (2000, 9000, 40000):
        seqs.append(random_seq(rng, L, star_frac=0.001, degen_frac=0.01))
        parts = []
        while sum(len(p) for p in parts) < L:
            parts.append(random_seq(rng, int(rng.integers(200, 2000)), 0.0, 0.0))
            parts.append(emit_homolog(orc, hmms[int(rng.integers(len(hmms)))], rng, int(rng.integers(60, 300))))
        seqs.append(np.concatenate(parts)[:L].astype(np.uint8))
    profs = [orc.configure(h) for h in hmms]
    want = orc.Batch(profs, seqs).msv16(nthreads=8).reshape(len(profs), len(seqs))
    gctx.set_profiles(hmms)
    gctx.set_targets(seqs)
    F = orc.Q_UNITS_PER_NAT
    for table in ([], [(64, 64, 1, 0), (512, 128, 4, 3)], [(512, 64, 8, 3)]):
        gctx.set_shapes(4, table)
        got, _ = gctx.score_all(4, 1)
        assert np.array_equal(got, want), (table, np.argwhere(got != want)[:5])
    gctx.set_shapes(4, [])
    for i, p in enumerate(profs):
        for j, d in enumerate(seqs):
            u = float(orc.msv(p, d))
            assert want[i, j] >= min(math.floor(F * u), q_saturated_floor(orc, len(d))), (i, j)
