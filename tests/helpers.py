"""Shared helpers for the test suite (random small models and sequences)."""
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
