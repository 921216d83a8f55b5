"""GPU parity on the five benchmark workloads (sentinel-x_amd/synth.py), at sizes the
oracle finishes in seconds, through the C ABI:

  single-M200  all of it: 1 profile x 2,000 proteins;
  marker-set   about 20 of the 400 profiles (M 80-1,500, those with most homologues and a
               spread of lengths) x all 4,000 proteins, and the whole search against the
               product's host tail fed with the oracle's special rows;
  pfam-scale   30 profiles spanning the lengths of a 200-profile draw x all 50,000 proteins;
  precision    10 profiles (M 1,500-2,500) x 10 proteins of 8,000-35,000 residues with
               low-complexity runs;
  mag-sweep    100 marker profiles x 5 whole genomes.

MSV and Viterbi are bit-exact, Forward and Backward within 1e-3 nats of the FP64 oracle,
and the cascade records equal the oracle's per-pair cascade (pipeline.rs:159-189) on every
pair, as the reference checks its own pipeline (search.rs:316-355).  The models' STATS
come from the oracle's scores, so the device and the oracle filter with the same lines.
"""
import math
import os

import numpy as np
import pytest

from helpers import calibrate_with_oracle, check_cascade_records, fwd_boundaries, oracle_hmms, spread

pytestmark = pytest.mark.gpu

NTH = min(16, os.cpu_count() or 1)
FWD_TOL = 1e-3
F = dict(f1=0.02, f2=1e-3, f3=1e-5)


def _synth():
    from sentinel_x_amd import synth
    return synth


def _load(ctx, wl_hmms, seqs):
    ctx.set_profiles(wl_hmms)
    ctx.set_targets(seqs)
    for st in range(6):
        ctx.set_shapes(st, [])


def _cascade_parity(orc, ctx, hmms, profs, seqs, forms=((1, False), (1, True), (0, False)), bounds=(True, False)):
    """Every form of the device cascade against the oracle's per-pair cascade."""
    stage_o, sc_o = orc.Batch(profs, seqs).cascade(**F, nthreads=NTH)
    nt = len(seqs)
    fb = fwd_boundaries(hmms, F["f3"])
    straddles = 0
    for form, graph in forms:
        for bound in bounds:
            st, rec = ctx.run_cascade(form, use_graph=graph, msv_bound=bound, **F)
            assert st.n_msv_pass == int((stage_o >= 1).sum())
            straddles += check_cascade_records(rec, stage_o, sc_o, nt, FWD_TOL, fb)
            assert st.n_fwd_pass == int((rec["stage"] == 4).sum())
    return stage_o, straddles


# ---------------------------------------------------------------------------
# single-M200: the whole workload
# ---------------------------------------------------------------------------

@pytest.fixture(scope="module")
def cfg0(orc):
    synth = _synth()
    wl = synth.make_workload(0)
    profs = calibrate_with_oracle(orc, synth, wl.hmms, NTH)
    return wl, profs


def test_cfg0_stages_on_every_pair(orc, gctx, cfg0):
    wl, profs = cfg0
    _load(gctx, wl.hmms, wl.seqs)
    b = orc.Batch(profs, wl.seqs)
    shape = (1, len(wl.seqs))
    want = {0: b.msv(nthreads=NTH).reshape(shape), 1: b.viterbi(nthreads=NTH).reshape(shape),
            4: b.msv16(nthreads=NTH).reshape(shape)}
    f64 = {2: b.forward64(nthreads=NTH).reshape(shape), 3: b.backward64(nthreads=NTH).reshape(shape)}
    for form in (0, 1):
        for st, w in want.items():
            got, _ = gctx.score_all(st, form)
            assert np.array_equal(got, w), (form, st, np.argwhere(got != w)[:5])
        for st, w in f64.items():
            got, _ = gctx.score_all(st, form)
            assert np.max(np.abs(got - w)) <= FWD_TOL, (form, st)


def test_cfg0_cascade_on_every_pair(orc, gctx, cfg0):
    wl, profs = cfg0
    _load(gctx, wl.hmms, wl.seqs)
    stage_o, _ = _cascade_parity(orc, gctx, wl.hmms, profs, wl.seqs)
    # the 5% homologues carry the survivors: every stage sees work
    assert (stage_o == 4).sum() >= 20 and (stage_o >= 1).sum() > (stage_o == 4).sum()


# ---------------------------------------------------------------------------
# marker-set: 20 profiles x all proteins, and the search
# ---------------------------------------------------------------------------

@pytest.fixture(scope="module")
def cfg1(orc):
    synth = _synth()
    wl = synth.make_workload(1)
    ms = np.array([h.m for h in wl.hmms])
    nhom = np.bincount(wl.homolog_of[wl.homolog_of >= 0], minlength=len(wl.hmms))
    pick = sorted(set(int(i) for i in np.argsort(-nhom, kind="stable")[:10]) | set(spread(ms, 12)))
    hmms = [wl.hmms[i] for i in pick]
    profs = calibrate_with_oracle(orc, synth, hmms, NTH)
    return hmms, profs, wl.seqs


def test_cfg1_stages_on_every_pair(orc, gctx, cfg1):
    hmms, profs, seqs = cfg1
    assert min(h.m for h in hmms) <= 120 and max(h.m for h in hmms) >= 1024  # short and long-model kernels
    _load(gctx, hmms, seqs)
    b = orc.Batch(profs, seqs)
    shape = (len(profs), len(seqs))
    for st, fn in ((0, b.msv), (1, b.viterbi)):
        got, _ = gctx.score_all(st, 1)
        w = fn(nthreads=NTH).reshape(shape)
        assert np.array_equal(got, w), (st, np.argwhere(got != w)[:5])
    for st, fn in ((2, b.forward64), (3, b.backward64)):
        got, _ = gctx.score_all(st, 1)
        err = np.abs(got - fn(nthreads=NTH).reshape(shape))
        assert float(err.max()) <= FWD_TOL, (st, float(err.max()))


def test_cfg1_cascade_on_every_pair(orc, gctx, cfg1):
    hmms, profs, seqs = cfg1
    _load(gctx, hmms, seqs)
    stage_o, _ = _cascade_parity(orc, gctx, hmms, profs, seqs)
    assert (stage_o == 4).sum() >= 10


@pytest.mark.parametrize("cutoff", ["evalue", "ga", "nc", "tc"])
def test_cfg1_search_matches_the_host_tail_on_oracle_rows(orc, sxa, gctx, cfg1, cutoff):
    """sx_search (device cascade, device Forward / Backward with posterior decoding and
    region finding, host tail from the regions on) against sx_search_host fed with the
    oracle's Forward / Backward special rows for the pairs the oracle's cascade passes
    (regions found on the host): the same hits, domain coordinates and reporting flags,
    scores within 0.01 bits (the device sums Forward in another order)."""
    hmms, profs, seqs = cfg1
    _load(gctx, hmms, seqs)
    kw = dict(cutoff=cutoff, seq_e=10.0, dom_e=10.0, filters=True, **F)
    hits, doms = gctx.search(**kw)
    stage_o, _ = orc.Batch(profs, seqs).cascade(**F, nthreads=NTH)
    idx = np.flatnonzero(stage_o == 4)
    pp, tt = idx // len(seqs), idx % len(seqs)
    fwd, bwd, fsp, bsp = [], [], [], []
    for p, t in zip(pp, tt):
        v, sp = orc.forward(profs[p], seqs[t], specials=True)
        fwd.append(v)
        fsp.append(sp)
        b, spb = orc.backward(profs[p], seqs[t])
        bwd.append(float(b))
        bsp.append(spb)
    oh = oracle_hmms(orc, hmms)
    hh, hd = sxa.search_host(oh, seqs, pp, tt, fwd, bwd, fsp, bsp, **kw)
    assert len(hits) >= 5
    assert np.array_equal(hits[["profile", "target", "qlen", "tlen", "ndom"]],
                          hh[["profile", "target", "qlen", "tlen", "ndom"]])
    assert np.max(np.abs(hits["score"] - hh["score"])) <= 0.01
    assert np.max(np.abs(hits["bias"] - hh["bias"])) <= 0.01
    assert np.allclose(hits["evalue"], hh["evalue"], rtol=0.01, atol=0)
    assert np.allclose(hits["nexpected"], hh["nexpected"], rtol=1e-3, atol=1e-4)
    coords = ["hmm_from", "hmm_to", "ali_from", "ali_to", "env_from", "env_to"]
    assert np.array_equal(doms[coords], hd[coords])
    assert np.max(np.abs(doms["score"] - hd["score"])) <= 0.01
    assert np.array_equal(doms["reported"], hd["reported"])


# ---------------------------------------------------------------------------
# pfam-scale: 30 profiles x all 50,000 proteins
# ---------------------------------------------------------------------------

@pytest.fixture(scope="module")
def cfg2(orc):
    synth = _synth()
    wl = synth.make_workload(2, profile_scale=0.01)  # a 200-profile draw of the length distribution
    ms = np.array([h.m for h in wl.hmms])
    hmms = [wl.hmms[i] for i in spread(ms, 30)]
    profs = calibrate_with_oracle(orc, synth, hmms, NTH)
    return hmms, profs, wl.seqs


def test_cfg2_cascade_on_every_pair(orc, gctx, cfg2):
    hmms, profs, seqs = cfg2
    assert len(seqs) == 50000
    _load(gctx, hmms, seqs)
    stage_o, _ = _cascade_parity(orc, gctx, hmms, profs, seqs, forms=((1, False), (1, True)))
    assert (stage_o >= 1).sum() > 1000


def test_cfg2_msv16_bound_over_every_pair(orc, gctx, cfg2):
    """The int16 bound is at least 2^22 times the f32 MSV (or above its saturation floor) on
    every pair, and equals the oracle's bound bit for bit on a third of the profiles."""
    from test_oracle import q_saturated_floor
    hmms, profs, seqs = cfg2
    _load(gctx, hmms, seqs)
    q, _ = gctx.score_all(4, 1)
    u, _ = gctx.score_all(0, 1)
    lens = np.array([len(s) for s in seqs])
    floors = {L: q_saturated_floor(orc, int(L)) for L in np.unique(lens)}
    fl = np.array([floors[L] for L in lens], np.float64)
    need = np.minimum(np.floor(orc.Q_UNITS_PER_NAT * u.astype(np.float64)), fl[None, :])
    assert np.all(q.astype(np.float64) >= need)
    sub = list(range(0, len(profs), 3))
    w = orc.Batch([profs[i] for i in sub], seqs).msv16(nthreads=NTH).reshape(len(sub), -1)
    assert np.array_equal(q[sub], w)


# ---------------------------------------------------------------------------
# precision stress: long models x long, low-complexity proteins
# ---------------------------------------------------------------------------

@pytest.fixture(scope="module")
def cfg3(orc):
    synth = _synth()
    wl = synth.make_workload(3, profile_scale=0.05, target_scale=0.005)
    profs = calibrate_with_oracle(orc, synth, wl.hmms, NTH, n=60)
    return wl.hmms, profs, wl.seqs


def test_cfg3_forward_backward_within_fp64_tolerance(orc, gctx, cfg3):
    hmms, profs, seqs = cfg3
    lens = [len(s) for s in seqs]
    assert len(profs) == 10 and len(seqs) == 10 and min(lens) >= 8000 and max(lens) > 20000
    _load(gctx, hmms, seqs)
    b = orc.Batch(profs, seqs)
    shape = (len(profs), len(seqs))
    f64 = b.forward64(nthreads=NTH).reshape(shape)
    b64 = b.backward64(nthreads=NTH).reshape(shape)
    for form in (0, 1):
        fwd, _ = gctx.score_all(2, form)
        bwd, _ = gctx.score_all(3, form)
        assert float(np.abs(fwd - f64).max()) <= FWD_TOL
        assert float(np.abs(bwd - b64).max()) <= FWD_TOL
    # the reference returns f32: its score gives the same GA / TC / NC decisions on the
    # Forward bit score as the FP64 one
    fwd, _ = gctx.score_all(2, 1)
    nulls = np.array([float(orc.null_score(L)) for L in lens])
    ln2 = math.log(2.0)
    bits_g = (fwd.astype(np.float32).astype(np.float64) - nulls) / ln2
    bits_r = (f64 - nulls) / ln2
    for cut in ("ga", "tc", "nc"):
        c = np.array([getattr(h, cut)[0] for h in hmms])[:, None]
        assert np.array_equal(bits_g >= c, bits_r >= c), cut


def test_cfg3_msv_and_viterbi_bit_exact(orc, gctx, cfg3):
    hmms, profs, seqs = cfg3
    _load(gctx, hmms, seqs)
    b = orc.Batch(profs, seqs)
    shape = (len(profs), len(seqs))
    for st, fn in ((0, b.msv), (1, b.viterbi)):
        got, _ = gctx.score_all(st, 1)
        assert np.array_equal(got, fn(nthreads=NTH).reshape(shape)), st


# ---------------------------------------------------------------------------
# mag-sweep: 100 markers x 5 whole genomes
# ---------------------------------------------------------------------------

@pytest.fixture(scope="module")
def cfg4(orc):
    synth = _synth()
    wl = synth.make_workload(4, profile_scale=0.02, n_genomes=5)
    profs = calibrate_with_oracle(orc, synth, wl.hmms, NTH)
    return wl, profs


def test_cfg4_cascade_on_every_pair(orc, gctx, cfg4):
    wl, profs = cfg4
    assert len(wl.hmms) == 100 and len(set(wl.genome_of.tolist())) == 5 and len(wl.seqs) >= 7000
    _load(gctx, wl.hmms, wl.seqs)
    stage_o, _ = _cascade_parity(orc, gctx, wl.hmms, profs, wl.seqs, forms=((1, False), (1, True)))
    assert (stage_o == 4).sum() >= 10
