"""GPU parity of the four DP stages against the C oracle, through the C ABI.

MSV and Viterbi must be bit-identical to the oracle (f32 max-plus, dp.rs:92-189);
Forward/Backward must lie within FWD_TOL nats of the oracle's FP64 twin
(dp.rs:229-392 evaluated in double over the same f32 parameters).
"""
import numpy as np
import pytest

from helpers import emit_homolog, random_hmm, random_seq

pytestmark = pytest.mark.gpu

FWD_TOL = 1e-3  # nats, absolute, against the FP64 oracle

MSV_SHAPES = [(q, g) for q in (4, 8, 16, 32, 64) for g in (1, 2, 4, 8, 16, 32, 64)] + \
    [(48, 1), (96, 1), (48, 2), (96, 2), (48, 4), (128, 1), (128, 2), (160, 1), (192, 1), (256, 1)] + \
    [(q, 1) for q in (40, 56, 72, 80, 112, 144, 176, 208, 224)] + [(q, 4) for q in (24, 40, 56, 72, 80)] + \
    [(q, 8) for q in (24, 40, 48, 56)] + [(q, 16) for q in (12, 20, 24, 40, 48)] + [(q, 32) for q in (24, 40, 48)]
DP3_SHAPES = [(q, g, h) for q, h in ((4, 1), (8, 1), (12, 1), (16, 1), (16, 0), (32, 0), (48, 0))
              for g in (1, 2, 4, 8, 16, 32, 64)] + [(64, 1, 0), (64, 2, 0), (64, 4, 0)]
EDGE_LENGTHS = [1, 2, 3, 15, 16, 17, 31, 32, 33, 47, 48, 64, 65]
LONG_M = 1024  # models longer than this use the fixed global-emission shape


# MSV with 16-byte-aligned lane blocks: (q, g, table mode 3 = 512 / 4 = 1024 threads)
MSV_ALN_SHAPES = [(q, 16, h) for q in (12, 16, 20, 24, 32, 40, 48) for h in (3, 4)] + [(64, 16, 3)] + \
    [(q, 32, h) for q in (16, 24, 32) for h in (3, 4)] + [(q, 8, h) for q in (16, 20, 24, 32, 40, 48) for h in (3, 4)]


def lds_ok(stage, q, g):
    w0 = g * (q + 1) + 2
    wid = w0 if w0 % 2 else w0 + 1
    rows = 25 if stage == 0 else 33
    return rows * wid * 4 <= 160 * 1024


def make_targets(orc, hmms, n, seed, lmax=300, homolog_frac=0.1):
    rng = np.random.default_rng(seed)
    seqs = [random_seq(rng, L, star_frac=0.01, degen_frac=0.02) for L in EDGE_LENGTHS]
    while len(seqs) < n:
        if hmms and rng.random() < homolog_frac:
            h = hmms[int(rng.integers(len(hmms)))]
            seqs.append(emit_homolog(orc, h, rng, int(rng.integers(20, lmax))))
        else:
            seqs.append(random_seq(rng, int(rng.integers(1, lmax)), star_frac=0.01, degen_frac=0.02))
    return seqs


@pytest.fixture(scope="module")
def mixed(orc, gctx):
    ms = [1, 2, 3, 5, 17, 40, 64, 100, 150, 229, 300, 512, 700, 1000]
    hmms = [random_hmm(orc, m, 500 + i) for i, m in enumerate(ms)]
    seqs = make_targets(orc, hmms, 60, 7)
    profs = [orc.configure(h) for h in hmms]
    batch = orc.Batch(profs, seqs)
    gctx.set_profiles(hmms)
    gctx.set_targets(seqs)
    for st in range(4):
        gctx.set_shapes(st, [])
    return hmms, profs, seqs, batch


def test_profile_tables_bit_identical_to_oracle(gctx, mixed):
    hmms, profs, _, _ = mixed
    for i, p in enumerate(profs):
        msc, tsc, mo, tp, tbm = gctx.profile_tables(i)
        for a, b in ((msc, p.msc), (tsc, p.tsc), (mo, p.mo), (tp, p.tp)):
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
        assert np.float32(tbm) == np.float32(p.msv_tbm)


@pytest.mark.parametrize("form", [0, 1], ids=["plain", "batched"])
def test_msv_all_pairs_bit_exact(gctx, mixed, form):
    hmms, profs, seqs, batch = mixed
    got, _ = gctx.score_all(0, form)
    want = batch.msv(nthreads=8).reshape(len(profs), len(seqs))
    assert np.array_equal(got, want), np.argwhere(got != want)[:5]


@pytest.mark.parametrize("form", [0, 1], ids=["plain", "batched"])
def test_viterbi_all_pairs_bit_exact(gctx, mixed, form):
    hmms, profs, seqs, batch = mixed
    got, _ = gctx.score_all(1, form)
    want = batch.viterbi(nthreads=8).reshape(len(profs), len(seqs))
    assert np.array_equal(got, want), np.argwhere(got != want)[:5]


@pytest.mark.parametrize("form", [0, 1], ids=["plain", "batched"])
def test_forward_all_pairs_within_tolerance(gctx, mixed, form):
    hmms, profs, seqs, batch = mixed
    got, _ = gctx.score_all(2, form)
    want = batch.forward64(nthreads=8).reshape(len(profs), len(seqs))
    assert np.array_equal(np.isinf(got), np.isinf(want))
    fin = np.isfinite(want)
    err = np.abs(got[fin] - want[fin])
    assert err.max() <= FWD_TOL, err.max()


@pytest.mark.parametrize("form", [0, 1], ids=["plain", "batched"])
def test_backward_all_pairs_within_tolerance(orc, gctx, mixed, form):
    hmms, profs, seqs, batch = mixed
    got, _ = gctx.score_all(3, form)
    f64 = batch.forward64(nthreads=8).reshape(len(profs), len(seqs))
    fin = np.isfinite(f64)
    assert np.all(np.isfinite(got[fin]))
    assert np.abs(got[fin] - f64[fin]).max() <= FWD_TOL
    # spot-check against the oracle's FP64 backward twin itself
    for i in range(0, len(profs), 3):
        for j in range(0, len(seqs), 7):
            b64 = orc.backward64(profs[i], seqs[j])
            assert abs(got[i, j] - b64) <= FWD_TOL


def test_pair_list_matches_all_pairs(gctx, mixed):
    hmms, profs, seqs, batch = mixed
    rng = np.random.default_rng(3)
    n = 300
    pi = rng.integers(0, len(profs), n).astype(np.uint32)
    ti = rng.integers(0, len(seqs), n).astype(np.uint32)
    want = batch.msv(nthreads=8).reshape(len(profs), len(seqs))[pi, ti]
    for form in (0, 1):
        got, _ = gctx.score_pairs(0, pi, ti, form=form)
        assert np.array_equal(got, want)
    wantv = batch.viterbi(nthreads=8).reshape(len(profs), len(seqs))[pi, ti]
    gotv, _ = gctx.score_pairs(1, pi, ti)
    assert np.array_equal(gotv, wantv)


def _spec_close(a, b, tol=2e-3):
    assert a.shape == b.shape
    ia, ib = np.isneginf(a), np.isneginf(b)
    assert np.array_equal(ia, ib), np.argwhere(ia != ib)[:5]
    f = ~ia
    return np.abs(a[f] - b[f]).max(initial=0.0) <= tol * max(1.0, np.abs(b[f]).max(initial=1.0) / 100)


def test_forward_backward_specials_match_oracle(orc, gctx, mixed):
    hmms, profs, seqs, _ = mixed
    rng = np.random.default_rng(5)
    pi = rng.integers(0, len(profs), 40).astype(np.uint32)
    ti = rng.integers(0, len(seqs), 40).astype(np.uint32)
    fsp, fsc = gctx.specials(2, pi, ti)
    bsp, bsc = gctx.specials(3, pi, ti)
    for k in range(len(pi)):
        p, d = profs[pi[k]], seqs[ti[k]]
        f, osp = orc.forward(p, d, specials=True)
        assert _spec_close(fsp[k], osp)
        assert abs(fsc[k] - orc.forward64(p, d)) <= FWD_TOL
        b, obsp = orc.backward(p, d)
        assert _spec_close(bsp[k], obsp)
        assert abs(bsc[k] - orc.backward64(p, d)) <= FWD_TOL + 1e-4 * abs(bsc[k])


def _shape_case(orc, cap, seed):
    ms = sorted({max(1, cap // 2 + 1), max(1, (3 * cap) // 4), cap})
    hmms = [random_hmm(orc, m, seed + i) for i, m in enumerate(ms)]
    seqs = make_targets(orc, hmms, 24, seed, lmax=160)
    return hmms, [orc.configure(h) for h in hmms], seqs


@pytest.mark.parametrize("q,g", MSV_SHAPES, ids=[f"Q{q}G{g}" for q, g in MSV_SHAPES])
def test_every_msv_shape_bit_exact(orc, gctx, q, g):
    cap = min(q * g, LONG_M)  # longer models always take the fixed long-model shape
    if not lds_ok(0, q, g):
        with pytest.raises(Exception):
            gctx.set_shapes(0, [(cap, q, g)])
        return
    hmms, profs, seqs = _shape_case(orc, cap, 1000 + q * 131 + g)
    gctx.set_profiles(hmms)
    gctx.set_targets(seqs)
    gctx.set_shapes(0, [(cap, q, g)])
    got, _ = gctx.score_all(0, 1)
    gctx.set_shapes(0, [])
    want = orc.Batch(profs, seqs).msv(nthreads=8).reshape(len(profs), len(seqs))
    assert np.array_equal(got, want)


@pytest.mark.parametrize("q,g,h", MSV_ALN_SHAPES, ids=[f"Q{q}G{g}M{h}" for q, g, h in MSV_ALN_SHAPES])
def test_aligned_msv_shapes_bit_exact(orc, gctx, q, g, h):
    """Aligned lane blocks, 512 / 1024-thread workgroups: enough targets for several
    tiles per profile, the last one partial."""
    cap = min(q * g, LONG_M)
    rng = np.random.default_rng(q * 7 + g + h)
    ms = sorted({cap // 2 + 1, (3 * cap) // 4, cap, int(rng.integers(cap // 2 + 1, cap + 1))})
    hmms = [random_hmm(orc, m, 3000 + q * 131 + g + i) for i, m in enumerate(ms)]
    seqs = make_targets(orc, hmms, 150, 3000 + q + g + h, lmax=200)
    profs = [orc.configure(x) for x in hmms]
    gctx.set_profiles(hmms)
    gctx.set_targets(seqs)
    gctx.set_shapes(0, [(cap, q, g, h)])
    got, _ = gctx.score_all(0, 1)
    pairs = (np.arange(40) % len(hmms), (np.arange(40) * 7) % len(seqs))
    gp, _ = gctx.score_pairs(0, *pairs)
    gctx.set_shapes(0, [])
    want = orc.Batch(profs, seqs).msv(nthreads=8).reshape(len(profs), len(seqs))
    assert np.array_equal(got, want)
    assert np.array_equal(gp, want[pairs])


def test_aligned_modes_are_msv_and_viterbi_only(gctx, orc):
    gctx.set_profiles([random_hmm(orc, 100, 5)])
    for stage in (2, 3):
        with pytest.raises(Exception):
            gctx.set_shapes(stage, [(256, 16, 16, 3)])
    with pytest.raises(Exception):
        gctx.set_shapes(0, [(256, 64, 4, 3)])  # needs G = 8 or G >= 16
    with pytest.raises(Exception):
        gctx.set_shapes(0, [(256, 16, 16, 2)])  # the long-model mode is fixed


@pytest.mark.parametrize("q,g,h", [(8, 8, 1), (16, 16, 1), (16, 16, 0), (12, 16, 3), (16, 8, 3), (24, 64, 2)],
                         ids=["Q8G8R", "Q16G16R", "Q16G16L", "Q12G16A", "Q16G8A", "Q24G64GEM"])
def test_viterbi_row_max_keeps_d_for_positive_delete_scores(orc, gctx, q, g, h):
    """A model whose M->D or D->D probabilities exceed 1 (parseable, though not a valid
    model) has positive delete scores: a D value can then beat every M value of its row, so
    Viterbi's xE must keep the D row.  Valid models take the M-only maximum; both must
    match the oracle bit for bit."""
    cap = q * g if h != 2 else 1400
    ms = [max(2, cap // 2 + 1), cap]
    hmms = [random_hmm(orc, m, 6000 + m) for m in ms]
    for x in hmms[:1]:
        t = np.array(x.t, np.float32)
        t[1:x.m:3, 2] = 20.0  # MD: ln 20 = +3.0
        t[2:x.m:2, 6] = 20.0  # DD
        x.t = t
    seqs = make_targets(orc, hmms, 60, 6000 + q + g, lmax=220)
    profs = [orc.configure(x) for x in hmms]
    gctx.set_profiles(hmms)
    gctx.set_targets(seqs)
    if h != 2:
        gctx.set_shapes(1, [(cap, q, g, h)])
    got, _ = gctx.score_all(1, 1)
    gctx.set_shapes(1, [])
    want = orc.Batch(profs, seqs).viterbi(nthreads=8).reshape(len(profs), len(seqs))
    assert np.array_equal(got, want)
    assert np.isfinite(want[0]).all()  # the planted model scores every target


VIT_ALN_SHAPES = [(8, 8), (12, 8), (16, 8), (8, 16), (12, 16), (16, 16), (24, 16), (12, 32), (16, 32), (24, 32),
                  (32, 32)]


@pytest.mark.parametrize("q,g", VIT_ALN_SHAPES, ids=[f"Q{q}G{g}" for q, g in VIT_ALN_SHAPES])
def test_aligned_viterbi_shapes_bit_exact(orc, gctx, q, g):
    """Viterbi over aligned lane blocks (transition rows staged shifted to the cell they
    feed), 512-thread workgroups: several tiles per profile, the last one partial; '*'
    residues (no insert emission) and degenerate codes among the targets."""
    cap = min(q * g, LONG_M)
    ms = sorted({cap // 2 + 1, (3 * cap) // 4, cap})
    hmms = [random_hmm(orc, m, 4000 + q * 131 + g + i) for i, m in enumerate(ms)]
    seqs = make_targets(orc, hmms, 150, 4000 + q + g, lmax=200)
    profs = [orc.configure(x) for x in hmms]
    gctx.set_profiles(hmms)
    gctx.set_targets(seqs)
    gctx.set_shapes(1, [(cap, q, g, 3)])
    got, _ = gctx.score_all(1, 1)
    gctx.set_shapes(1, [])
    want = orc.Batch(profs, seqs).viterbi(nthreads=8).reshape(len(profs), len(seqs))
    assert np.array_equal(got, want)


@pytest.mark.parametrize("stage", [1, 2, 3], ids=["viterbi", "forward", "backward"])
@pytest.mark.parametrize("q,g,h", DP3_SHAPES, ids=[f"Q{q}G{g}{'R' if h else 'L'}" for q, g, h in DP3_SHAPES])
def test_every_dp3_shape(orc, gctx, q, g, h, stage):
    cap = min(q * g, LONG_M)
    if not lds_ok(stage, q, g):
        with pytest.raises(Exception):
            gctx.set_shapes(stage, [(cap, q, g, h)])
        return
    hmms, profs, seqs = _shape_case(orc, cap, 2000 + q * 131 + g)
    gctx.set_profiles(hmms)
    gctx.set_targets(seqs)
    gctx.set_shapes(stage, [(cap, q, g, h)])
    got, _ = gctx.score_all(stage, 1)
    gctx.set_shapes(stage, [])
    batch = orc.Batch(profs, seqs)
    if stage == 1:
        want = batch.viterbi(nthreads=8).reshape(len(profs), len(seqs))
        assert np.array_equal(got, want)
    else:
        want = batch.forward64(nthreads=8).reshape(len(profs), len(seqs))
        fin = np.isfinite(want)
        assert np.array_equal(np.isfinite(got), fin)
        assert np.abs(got[fin] - want[fin]).max() <= FWD_TOL


VIT_FINE_SHAPES = [(q, g) for q in (10, 14) for g in (4, 8, 16, 32)]


@pytest.mark.parametrize("q,g", VIT_FINE_SHAPES, ids=[f"Q{q}G{g}" for q, g in VIT_FINE_SHAPES])
def test_fine_viterbi_shapes_bit_exact(orc, gctx, q, g):
    """The Viterbi-only shapes with Q = 10 and 14 positions a lane (transitions in
    registers): bit-exact with the oracle, models up to their capacity."""
    cap = q * g
    hmms, profs, seqs = _shape_case(orc, cap, 2600 + q * 131 + g)
    gctx.set_profiles(hmms)
    gctx.set_targets(seqs)
    gctx.set_shapes(1, [(cap, q, g, 1)])
    got, _ = gctx.score_all(1, 1)
    gctx.set_shapes(1, [])
    want = orc.Batch(profs, seqs).viterbi(nthreads=8).reshape(len(profs), len(seqs))
    assert np.array_equal(got, want)


def test_empty_target_scores_neg_inf(orc, gctx):
    hmms = [random_hmm(orc, 20, 77)]
    gctx.set_profiles(hmms)
    gctx.set_targets([np.zeros(0, np.uint8), np.array([3, 4, 5], np.uint8)])
    for st in range(4):
        got, _ = gctx.score_all(st, 1)
        assert np.isneginf(got[0, 0])
        assert np.isfinite(got[0, 1])


def test_autotuned_tables_keep_parity(orc, gctx):
    ms = [20, 60, 150, 150, 300, 700]
    hmms = [random_hmm(orc, m, 900 + i) for i, m in enumerate(ms)]
    seqs = make_targets(orc, hmms, 200, 9)
    profs = [orc.configure(h) for h in hmms]
    gctx.set_profiles(hmms)
    gctx.set_targets(seqs)
    batch = orc.Batch(profs, seqs)
    for st in (0, 1, 2):
        gctx.autotune(st, 500.0)
        table = gctx.get_shapes(st)
        assert table and table[-1][0] >= max(ms)
        for mmax, q, g, h in table:
            assert q * g >= mmax
    shape = (len(profs), len(seqs))
    assert np.array_equal(gctx.score_all(0, 1)[0], batch.msv(nthreads=8).reshape(shape))
    assert np.array_equal(gctx.score_all(1, 1)[0], batch.viterbi(nthreads=8).reshape(shape))
    f = gctx.score_all(2, 1)[0]
    assert np.abs(f - batch.forward64(nthreads=8).reshape(shape)).max() <= FWD_TOL
    for st in range(4):
        gctx.set_shapes(st, [])
