"""Shared helpers for the test suite (random small models and sequences)."""
import math

import numpy as np


def random_hmm(orc, m, seed, peaked=True, stats=True):
    """A random Plan7 model in probability space, shaped like parse.rs output."""
    rng = np.random.default_rng(seed)
    bg = orc.background().astype(np.float64)
    mat = np.zeros((m + 1, 20), np.float32)
    alpha = 0.5 * bg * 20 if peaked else bg * 50
    for k in range(1, m + 1):
        mat[k] = rng.dirichlet(alpha).astype(np.float32)
    t = np.zeros((m + 1, 7), np.float32)
    for k in range(0, m + 1):
        mm, mi, md = rng.dirichlet([18, 1, 1])
        im, ii = rng.dirichlet([7, 3])
        dm, dd = rng.dirichlet([7, 3])
        t[k] = [mm, mi, md, im, ii, dm, dd]
    t[m] = [1.0, 0.0, 0.0, 1.0, 0.0, 1.0, 0.0]
    t[0, 5], t[0, 6] = 1.0, 0.0
    compo = mat[1:].mean(axis=0).astype(np.float32)
    h = orc.Hmm(name=f"rand{seed}", m=m, mat=mat, t=t, compo=compo)
    if stats:
        h.stats = {"MSV": (-9.0, 0.69), "VITERBI": (-10.0, 0.69), "FORWARD": (-4.0, 0.69)}
    return h


def random_seq(rng, L, star_frac=0.0, degen_frac=0.0):
    s = rng.integers(0, 20, size=L).astype(np.uint8)
    if degen_frac:
        mask = rng.random(L) < degen_frac
        s[mask] = rng.integers(20, 24, size=int(mask.sum()))
    if star_frac:
        mask = rng.random(L) < star_frac
        s[mask] = 24
    return s


def emit_homolog(orc, h, rng, L):
    """Sample a local path through the model (entry uniform over k) embedded in
    background flanks; returns a code sequence of length about L."""
    bg = orc.background().astype(np.float64)
    bg = bg / bg.sum()
    m = h.m
    k = int(rng.integers(1, max(2, m // 3) + 1))
    core = []
    while k <= m and len(core) < L:
        core.append(int(rng.choice(20, p=h.mat[k] / h.mat[k].sum())))
        tk = h.t[k].astype(np.float64)
        r = rng.random()
        if r < tk[1] / tk[:3].sum():
            core.append(int(rng.choice(20, p=bg)))
        k += 1
    core = core[:L]
    left = int(rng.integers(0, max(1, L - len(core)) + 1))
    right = max(0, L - len(core) - left)
    s = list(rng.choice(20, size=left, p=bg)) + core + list(rng.choice(20, size=right, p=bg))
    return np.array(s, np.uint8)


def oracle_hmms(orc, hmms):
    """Synthetic workload models (synth.SynthHmm) as the oracle's Hmm records."""
    return [orc.Hmm(name=h.name, m=h.m, mat=h.mat, t=h.t, compo=h.compo, acc=h.acc, desc=h.desc, ga=h.ga, tc=h.tc,
                    nc=h.nc, stats=h.stats) for h in hmms]


def calibrate_with_oracle(orc, synth, hmms, nthreads, n=200):
    """STATS lines for every model (synth.calibrate), scored by the oracle."""
    profs = [orc.configure(h) for h in oracle_hmms(orc, hmms)]

    def scorer(stage, ss):
        b = orc.Batch(profs, ss)
        fn = {0: b.msv, 1: b.viterbi, 2: b.forward}[stage]
        return fn(nthreads=nthreads).reshape(len(profs), len(ss))

    synth.calibrate(hmms, scorer, n=n)
    return [orc.configure(h) for h in oracle_hmms(orc, hmms)]


def check_cascade_records(rec, stage_o, sc_o, nt, fwd_tol=1e-3, fwd_bounds=None):
    """Device cascade records against the oracle's per-pair cascade (Batch.cascade: stage
    and (msv, bias, viterbi, forward) per flat pair p * nt + t): the same MSV survivors,
    MSV, bias and Viterbi scores equal bit for bit, Forward within fwd_tol nats of the
    oracle's f32 Forward, and equal stages.  The Forward scores are summed in another order
    than the oracle's, so a pair whose two Forward bit scores straddle its F3 boundary
    (fwd_bounds[p], in bits) may end at stage 3 on one side and 4 on the other; such pairs
    are counted and returned, every other disagreement fails."""
    key = rec["profile"].astype(np.int64) * nt + rec["target"].astype(np.int64)
    order = np.argsort(key)
    rec, key = rec[order], key[order]
    assert np.array_equal(key, np.flatnonzero(stage_o >= 1)), "MSV survivor sets differ"
    so, sc = stage_o[key], sc_o[key]
    assert np.array_equal(rec["msv"], sc[:, 0]), "MSV scores differ"
    assert np.array_equal(rec["bias"], sc[:, 1]), "bias filter scores differ"
    reached = so >= 2
    assert np.array_equal(rec["vit"][reached], sc[reached, 2]), "Viterbi scores differ"
    f3 = so >= 3
    assert np.all(np.isnan(rec["fwd"][~(rec["stage"] >= 3)]))
    both = f3 & (rec["stage"] >= 3)
    assert np.all(np.abs(rec["fwd"][both].astype(np.float64) - sc[both, 3]) <= fwd_tol)
    diff = np.flatnonzero(rec["stage"] != so)
    straddle = 0
    for i in diff:
        assert {int(rec["stage"][i]), int(so[i])} == {3, 4}, (int(key[i]), int(rec["stage"][i]), int(so[i]))
        ln2 = np.float32(math.log(2.0))
        xd = float((np.float32(rec["fwd"][i]) - np.float32(rec["bias"][i])) / ln2)
        xo = float((np.float32(sc[i, 3]) - np.float32(sc[i, 1])) / ln2)
        xs = fwd_bounds[int(key[i]) // nt]
        assert min(xd, xo) <= xs + 1e-6 and max(xd, xo) >= xs - 1e-6, "Forward decision differs off its boundary"
        straddle += 1
    return straddle


def fwd_boundaries(hmms, f3):
    """Per profile, the Forward bit score at which exp(exp_logsurv) equals F3."""
    return np.array([h.stats["FORWARD"][0] - math.log(f3) / h.stats["FORWARD"][1] for h in hmms])


def spread(ms, k):
    """k indices whose values of ms cover their range (sorted order, evenly spaced)."""
    order = np.argsort(ms, kind="stable")
    return sorted({int(order[int(round(i * (len(ms) - 1) / max(1, k - 1)))]) for i in range(k)})
