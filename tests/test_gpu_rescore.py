"""Envelope rescoring on the device (sx_rescore: rescore_isolated_domain,
pipeline.rs:545-591) against two checks:

* the host restatement of the same steps (sx_rescore_host: forward_matrix,
  backward_matrix, posterior decoding, null2 and optimal_accuracy_coords in the
  reference's operation order, dp.rs:442-601 and 1274-1564; pinned end to end by the
  HMMER 3.4 fixture): the same status, the same optimal-accuracy coordinates, null2 odds
  within f32 summation-order rounding;
* the oracle with the unihit length model of the whole target (model.rs:284-312 with
  multihit = false): the envelope's Forward and Backward scores within 1e-3 nats of the
  oracle's f32 restatements of dp.rs:229-298 and 301-392, which the host restatement
  equals bit for bit.  Not the FP64 twins: on unihit envelopes the reference's own f32
  passes can sit nats below FP64 (after a row rescaling by a large E the N-state mass
  underflows, and later alignment starts drop out), and the scores and posteriors the
  reference goes on with are those f32 values.

Envelopes: the domains of the HMMER 3.4 fixture pairs and of a synthetic marker-set slice
with planted homologues (models of 80-1,500 positions, so the long-model widths run too),
widened and shifted copies of them, whole targets, envelopes of one and two rows, and a
homologue with '*' residues inside its envelope.  Then the search itself: device and host
rescoring give the same hits and coordinates."""
import os

import numpy as np
import pytest

from test_search_host import GOLDEN, read_faa

pytestmark = pytest.mark.gpu

# the synthetic models carry no STATS lines (no E-values), so hits are taken by bit score
BITS = dict(cutoff="bits", seq_t=5.0, dom_t=0.0)


@pytest.fixture(scope="module")
def fixture34(orc):
    hmms = orc.parse_hmm_file(os.path.join(GOLDEN, "hmmer34_pair.hmm"))
    tnames, raw = read_faa(os.path.join(GOLDEN, "hmmer34_targets.faa"))
    return hmms, [orc.encode(s) for s in raw]


@pytest.fixture(scope="module")
def marker(orc):
    """13 synthetic models spread over 80-2,200 positions and 300 proteins, 20% of them
    emitted from the models (with flanks), two with '*' residues planted in the homologue.
    Models up to 1,995 positions run the kernel form with the transition rows in LDS, the
    2,200-position model the form that reads them through the caches."""
    from conftest import load_package
    load_package()
    from sentinel_x_amd import synth
    rng = np.random.default_rng(2024)
    ms = [80, 120, 180, 250, 330, 420, 520, 700, 900, 1100, 1300, 1500, 2200]
    hmms = [synth.make_profile(rng, m, f"env{m}") for m in ms]
    lens = synth._lognormal_int(rng, 280, 0.5, 50, 2000, 300)
    # homologues are emitted from a local path with flanks, so lengthen those of the
    # long models' homologues to hold them
    seqs, _ = synth._proteins(rng, np.maximum(lens, 400), hmms, 0.2)
    return hmms, [np.array(s, np.uint8) for s in seqs]


def envelope_set(hits, doms, seqs, rng, n_random=40):
    """(profile, target, i, j, null2) rows built from a search's domains."""
    env = []
    for h in hits:
        p, t = int(h["profile"]), int(h["target"])
        L = len(seqs[t])
        for d in doms[int(h["dom_first"]):int(h["dom_first"]) + int(h["ndom"])]:
            i, j = int(d["env_from"]), int(d["env_to"])
            env.append((p, t, i, j, 1))
            env.append((p, t, max(1, i - 7), min(L, j + 5), 0))
            env.append((p, t, min(L, i + 3), max(min(L, i + 3), j - 3), 1))
        env.append((p, t, 1, L, 1))
        env.append((p, t, 1, 1, 1))
        env.append((p, t, L, L, 0))
        env.append((p, t, max(1, L // 2), min(L, L // 2 + 1), 1))
        for _ in range(n_random // max(1, len(hits))):
            a, b = sorted(rng.integers(1, L + 1, 2))
            env.append((p, t, int(a), int(b), int(rng.integers(0, 2))))
    return env


def check(orc, sxa, gctx, profs, seqs, env, nat_tol=1e-3):
    dev = gctx.rescore(env)
    host = gctx.rescore(env, on_host=True)
    assert np.array_equal(dev["status"], host["status"])
    ok = dev["status"] == 0
    assert ok.sum() >= len(env) // 2
    assert np.array_equal(dev["oa"], host["oa"])
    for k in ("iali", "jali", "kali", "mali"):
        assert np.array_equal(dev[k], host[k]), k
    e = np.asarray(env, sxa.ENVELOPE_DTYPE)
    n2 = ok & (e["null2"] == 1)
    assert np.allclose(dev["null2"][n2], host["null2"][n2], rtol=2e-4, atol=1e-6)
    # the envelope scores against the oracle with the whole target's unihit length model
    for r, (p, t, i, j, _) in enumerate(env):
        if not ok[r]:
            continue
        xs = orc.xsc(len(seqs[t]), multihit=False)
        sub = seqs[t][i - 1:j]
        f32 = float(np.float32(orc.forward(profs[p], sub, xs=xs)))
        b32 = float(orc.backward(profs[p], sub, xs=xs)[0])
        assert abs(float(dev["envsc"][r]) - f32) <= nat_tol, (r, p, t, i, j)
        assert abs(float(dev["bcksc"][r]) - b32) <= nat_tol, (r, p, t, i, j)
        assert float(host["envsc"][r]) == f32 and float(host["bcksc"][r]) == b32
    return dev


def test_rescore_fixture_envelopes(orc, sxa, gctx, fixture34):
    hmms, seqs = fixture34
    gctx.set_profiles(hmms)
    gctx.set_targets(seqs)
    hits, doms = gctx.search(filters=False, host_rescore=True)
    env = envelope_set(hits, doms, seqs, np.random.default_rng(7))
    assert len(env) > 100
    dev = check(orc, sxa, gctx, [orc.configure(h) for h in hmms], seqs, env)
    assert dev["oa"].sum() >= len(hits)


def test_rescore_marker_envelopes_long_models_and_stars(orc, sxa, gctx, marker):
    hmms, seqs = marker
    seqs = [s.copy() for s in seqs]
    gctx.set_profiles(hmms)
    gctx.set_targets(seqs)
    hits, doms = gctx.search(filters=False, host_rescore=True, **BITS)
    assert len(hits) >= 10
    # '*' residues inside two homologues' first domains
    starred = []
    for h in hits[:2]:
        d = doms[int(h["dom_first"])]
        t = int(h["target"])
        mid = (int(d["env_from"]) + int(d["env_to"])) // 2
        seqs[t][mid - 1] = 24
        seqs[t][min(len(seqs[t]), mid + 3) - 1] = 24
        starred.append(t)
    gctx.set_targets(seqs)
    env = envelope_set(hits, doms, seqs, np.random.default_rng(11), n_random=60)
    ms = {int(hits["profile"][i]): hmms[int(hits["profile"][i])].m for i in range(len(hits))}
    assert max(ms.values()) > 1995, "the slice has no hit of the longest model"
    check(orc, sxa, gctx, [orc.configure(h) for h in hmms], seqs, env)


def test_rescore_rejects_envelopes_outside_the_target(sxa, gctx, fixture34):
    hmms, seqs = fixture34
    gctx.set_profiles(hmms)
    gctx.set_targets(seqs)
    L = len(seqs[0])
    for bad in [(0, 0, 0, 3, 1), (0, 0, 5, 4, 1), (0, 0, 1, L + 1, 1), (0, len(seqs), 1, 1, 1), (len(hmms), 0, 1, 2, 1)]:
        with pytest.raises(RuntimeError):
            gctx.rescore([bad])
    assert len(gctx.rescore([])) == 0


@pytest.mark.parametrize("filters", [False, True])
def test_search_with_device_rescoring_matches_host_rescoring(gctx, marker, filters):
    """sx_search with the envelopes rescored on the device and on host threads: the same
    hits, domains, coordinates and reporting flags; scores within 1e-3 bits plus 5e-5 of
    the score (the envelope sums over up to 1,500 x 2,000 cells run in another order, so
    a 1,000-bit hit moves by f32 rounding of its sums: 1e-5 relative is typical)."""
    hmms, seqs = marker
    gctx.set_profiles(hmms)
    gctx.set_targets(seqs)
    hd, dd = gctx.search(filters=filters, **BITS)
    st = dict(gctx.last_search_stats)
    hh, dh = gctx.search(filters=filters, host_rescore=True, **BITS)
    sh = dict(gctx.last_search_stats)
    assert len(hd) >= 10
    assert st["n_envelopes"] >= len(dd) and st["ms_rescore"] > 0.0
    assert sh["n_envelopes"] == 0 and sh["ms_rescore"] == 0.0
    assert np.array_equal(hd[["profile", "target", "ndom", "qlen", "tlen"]], hh[["profile", "target", "ndom", "qlen", "tlen"]])
    assert np.allclose(hd["score"], hh["score"], rtol=5e-5, atol=1e-3)
    # bias = pre-null2 score - score: the difference of two such sums, so its rounding
    # scales with the score
    assert np.all(np.abs(hd["bias"] - hh["bias"]) <= 1e-3 + 1e-4 * np.abs(hh["score"]))
    for k in ("hmm_from", "hmm_to", "ali_from", "ali_to", "env_from", "env_to", "reported"):
        assert np.array_equal(dd[k], dh[k]), k
    assert np.allclose(dd["score"], dh["score"], rtol=5e-5, atol=1e-3)
