"""Synthetic Plan7 profiles and protein sets for the five benchmark workloads.

Workloads (seeded, numpy PCG64):
  0 "single-M200":   1 profile, M=200, vs 2,000 proteins (L ~ lognormal(median 300, s=0.5) in [50, 2000]),
                     5% emitted from the profile.
  1 "marker-set":    400 profiles, M ~ lognormal(median 350, s=0.5) in [80, 1500], vs 4,000 proteins
                     (median 280, s=0.5, [50, 2000]); 3% homologues of random profiles.
  2 "pfam-scale":    20,000 profiles, M ~ lognormal(median 150, s=0.6) in [25, 1000], vs 50,000 proteins
                     (median 300, s=0.5, [50, 5000]).
  3 "precision":     200 profiles, M ~ U[1500, 2500], vs 2,000 proteins, L ~ U[8000, 35000], 20% of the
                     residues in runs drawn from {K, Q, E, S}.
  4 "mag-sweep":     5,000 profiles (median 300, s=0.5, [25, 2500]) vs 200 genomes of U[1.5, 6] Mbp with
                     one protein per kb (median 280, s=0.5), 2% homologues.

Profiles: match emissions ~ Dirichlet(0.5 * background), floored at 1e-5 and renormalised (HMMER's
priors never leave a zero emission, and a zero would make the match score -inf); insert emissions =
background; (MM, MI, MD) ~ Dir(18, 1, 1), (IM, II) ~ Dir(7, 3), (DM, DD) ~ Dir(7, 3); node M is
"MM=1, IM=1, DM=1"; COMPO = mean match emission; GA/TC/NC = 25/20, 30/25, 20/15 bits.
`calibrate` fits STATS the way HMMER's calibration does (MSV and Viterbi Gumbel location for 200
background sequences of L=200 with lambda fixed; Forward exponential tail from L=100, tail mass 0.04)
using whatever scorer it is given (the device kernels in the benchmark).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Callable, List, Optional

import numpy as np

BG = np.array([0.0787945, 0.0151600, 0.0535222, 0.0668298, 0.0397062, 0.0695071, 0.0229198, 0.0590092,
               0.0594422, 0.0963728, 0.0237718, 0.0414386, 0.0482904, 0.0395639, 0.0540978, 0.0683364,
               0.0540687, 0.0673417, 0.0114135, 0.0304133], np.float64)
BGN = BG / BG.sum()
LAMBDA = 0.693147  # Gumbel / exponential-tail slope (bits), as HMMER's calibrated models carry
EMIT_FLOOR = 1e-5
KQES = np.array([8, 13, 3, 15])  # K Q E S


@dataclass
class SynthHmm:
    name: str
    m: int
    mat: np.ndarray
    t: np.ndarray
    compo: np.ndarray
    acc: Optional[str] = None
    desc: Optional[str] = None
    ga: tuple = (25.0, 20.0)
    tc: tuple = (30.0, 25.0)
    nc: tuple = (20.0, 15.0)
    stats: dict = field(default_factory=dict)


@dataclass
class Workload:
    name: str
    hmms: List[SynthHmm]
    seqs: List[np.ndarray]
    genome_of: Optional[np.ndarray] = None  # workload 4: genome index of each protein
    homolog_of: Optional[np.ndarray] = None  # profile index a protein was emitted from, or -1

    @property
    def cells(self) -> int:
        ms = np.array([h.m for h in self.hmms], np.int64)
        ls = np.array([len(s) for s in self.seqs], np.int64)
        return int(ms.sum() * ls.sum())


def _dirichlet_log(rng, alpha, n):
    """n draws of Dirichlet(alpha) computed in log space (stable for alpha << 1)."""
    a = np.asarray(alpha, np.float64)
    # X = Gamma(a+1) * U^(1/a)  =>  log X = log Gamma(a+1) + log(U) / a
    lg = np.log(rng.gamma(a + 1.0, size=(n, a.size))) + np.log(rng.random((n, a.size))) / a
    lg -= lg.max(axis=1, keepdims=True)
    p = np.exp(lg)
    return p / p.sum(axis=1, keepdims=True)


def make_profile(rng, m: int, name: str) -> SynthHmm:
    mat = np.zeros((m + 1, 20), np.float64)
    e = _dirichlet_log(rng, 0.5 * BGN, m)
    e = np.maximum(e, EMIT_FLOOR)
    mat[1:] = e / e.sum(axis=1, keepdims=True)
    t = np.zeros((m + 1, 7), np.float64)
    t[:, 0:3] = rng.dirichlet([18.0, 1.0, 1.0], size=m + 1)
    t[:, 3:5] = rng.dirichlet([7.0, 3.0], size=m + 1)
    t[:, 5:7] = rng.dirichlet([7.0, 3.0], size=m + 1)
    t[0, 5], t[0, 6] = 1.0, 0.0
    t[m] = [1.0, 0.0, 0.0, 1.0, 0.0, 1.0, 0.0]
    compo = mat[1:].mean(axis=0)
    return SynthHmm(name=name, m=m, mat=mat.astype(np.float32), t=t.astype(np.float32),
                    compo=compo.astype(np.float32))


def _lognormal_int(rng, median, sigma, lo, hi, n):
    v = np.exp(rng.normal(math.log(median), sigma, n))
    return np.clip(np.rint(v), lo, hi).astype(np.int64)


def background_seq(rng, L: int) -> np.ndarray:
    return rng.choice(20, size=L, p=BGN).astype(np.uint8)


def emit_homolog(rng, h: SynthHmm, L: int) -> np.ndarray:
    """A local path through the model (entry in the first third, exit anywhere past it),
    Plan7 transitions, match emissions, background inserts, embedded in background flanks."""
    m = h.m
    k = int(rng.integers(1, max(1, m // 3) + 1))
    kend = int(rng.integers(min(m, k + max(1, m // 3)), m + 1))
    core: List[int] = []
    state = "M"
    limit = 4 * L
    while k <= kend and len(core) < limit:
        tk = h.t[k].astype(np.float64)
        if state == "M":
            core.append(int(rng.choice(20, p=h.mat[k] / h.mat[k].sum())))
            r = rng.random() * tk[0:3].sum()
            state = "M" if r < tk[0] else ("I" if r < tk[0] + tk[1] else "D")
            if state != "I":
                k += 1
        elif state == "I":
            core.append(int(rng.choice(20, p=BGN)))
            state = "I" if rng.random() * tk[3:5].sum() >= tk[3] else "M"
            if state == "M":
                k += 1
        else:
            state = "D" if rng.random() * tk[5:7].sum() >= tk[5] else "M"
            k += 1
    core_a = np.array(core[:max(1, L)], np.uint8)
    if len(core_a) >= L:
        return core_a[:L]
    left = int(rng.integers(0, L - len(core_a) + 1))
    right = L - len(core_a) - left
    return np.concatenate([background_seq(rng, left), core_a, background_seq(rng, right)]).astype(np.uint8)


def _proteins(rng, lens, hmms, frac):
    seqs, hom = [], np.full(len(lens), -1, np.int64)
    for i, L in enumerate(lens):
        if hmms and rng.random() < frac:
            p = int(rng.integers(len(hmms)))
            hom[i] = p
            seqs.append(emit_homolog(rng, hmms[p], int(L)))
        else:
            seqs.append(background_seq(rng, int(L)))
    return seqs, hom


def _low_complexity(rng, L):
    """20% of residues in runs drawn from a 4-letter subset."""
    s = background_seq(rng, L)
    target = int(0.2 * L)
    done = 0
    while done < target:
        run = int(rng.integers(10, 60))
        start = int(rng.integers(0, max(1, L - run)))
        s[start:start + run] = rng.choice(KQES, size=min(run, L - start))
        done += run
    return s


WORKLOAD_NAMES = {0: "single-M200", 1: "marker-set", 2: "pfam-scale", 3: "precision", 4: "mag-sweep"}


def make_workload(k: int, profile_scale: float = 1.0, target_scale: float = 1.0,
                  n_genomes: Optional[int] = None) -> Workload:
    """Build workload k; the scale factors shrink the profile / protein counts (for tests).
    For workload 4, n_genomes (if given) sets the number of genomes, each of full size
    (target_scale then shrinks only the genome count when n_genomes is None)."""
    ss = np.random.SeedSequence(1000 + k)
    r_prof, r_len, r_res, r_hom = [np.random.default_rng(s) for s in ss.spawn(4)]

    def npr(n):
        return max(1, int(round(n * profile_scale)))

    def ntg(n):
        return max(1, int(round(n * target_scale)))

    genome_of = None
    if k == 0:
        ms = np.array([200])
        hmms = [make_profile(r_prof, 200, "synth0_0")]
        lens = _lognormal_int(r_len, 300, 0.5, 50, 2000, ntg(2000))
        seqs, hom = _proteins(r_res, lens, hmms, 0.05)
    elif k == 1:
        ms = _lognormal_int(r_len, 350, 0.5, 80, 1500, npr(400))
        hmms = [make_profile(r_prof, int(m), f"synth1_{i}") for i, m in enumerate(ms)]
        lens = _lognormal_int(r_len, 280, 0.5, 50, 2000, ntg(4000))
        seqs, hom = _proteins(r_res, lens, hmms, 0.03)
    elif k == 2:
        ms = _lognormal_int(r_len, 150, 0.6, 25, 1000, npr(20000))
        hmms = [make_profile(r_prof, int(m), f"synth2_{i}") for i, m in enumerate(ms)]
        lens = _lognormal_int(r_len, 300, 0.5, 50, 5000, ntg(50000))
        seqs, hom = _proteins(r_res, lens, hmms, 0.0)
    elif k == 3:
        ms = r_len.integers(1500, 2501, npr(200))
        hmms = [make_profile(r_prof, int(m), f"synth3_{i}") for i, m in enumerate(ms)]
        lens = r_len.integers(8000, 35001, ntg(2000))
        seqs = [_low_complexity(r_res, int(L)) for L in lens]
        hom = np.full(len(lens), -1, np.int64)
    elif k == 4:
        ms = _lognormal_int(r_len, 300, 0.5, 25, 2500, npr(5000))
        hmms = [make_profile(r_prof, int(m), f"synth4_{i}") for i, m in enumerate(ms)]
        ngen = n_genomes if n_genomes is not None else ntg(200)
        sizes_kb = r_len.uniform(1500, 6000, ngen)
        per_kb = 1.0 if n_genomes is not None else target_scale
        counts = np.maximum(1, np.rint(sizes_kb * per_kb).astype(np.int64))
        genome_of = np.repeat(np.arange(ngen), counts)
        lens = _lognormal_int(r_len, 280, 0.5, 30, 5000, int(counts.sum()))
        seqs, hom = _proteins(r_res, lens, hmms, 0.02)
    else:
        raise ValueError(f"unknown workload {k}")
    return Workload(WORKLOAD_NAMES[k], hmms, seqs, genome_of, hom)


# ---------------------------------------------------------------------------
# calibration (HMMER-style STATS lines)
# ---------------------------------------------------------------------------

def null_score(L):
    p1 = np.float32(L) / (np.float32(L) + np.float32(1.0))
    return np.float32(L * math.log(float(p1)) + math.log(1.0 - float(p1)))


def gumbel_mu(bits: np.ndarray, lam: float) -> float:
    """ML Gumbel location with known slope: mu = -(1/lam) log(mean(exp(-lam x)))."""
    x = np.asarray(bits, np.float64)
    return float(-np.log(np.mean(np.exp(-lam * (x - x.min())))) / lam + x.min())


def calibrate(hmms: List[SynthHmm], scorer: Callable, n: int = 200, seed: int = 4242):
    """scorer(stage, seqs) -> scores array (P, n) in nats for stage 0 (MSV), 1 (Viterbi), 2 (Forward)
    over the given hmms.  Sets each hmm's stats in place."""
    rng = np.random.default_rng(seed)
    s200 = [background_seq(rng, 200) for _ in range(n)]
    s100 = [background_seq(rng, 100) for _ in range(n)]
    ln2 = math.log(2.0)
    msv = scorer(0, s200)
    vit = scorer(1, s200)
    fwd = scorer(2, s100)
    n200, n100 = float(null_score(200)), float(null_score(100))
    for i, h in enumerate(hmms):
        mmu = gumbel_mu((msv[i] - n200) / ln2, LAMBDA)
        vmu = gumbel_mu((vit[i] - n200) / ln2, LAMBDA)
        fmu = gumbel_mu((np.asarray(fwd[i], np.float64) - n100) / ln2, LAMBDA)
        tailp = 0.04
        tau = fmu - math.log(-math.log(1.0 - tailp)) / LAMBDA + math.log(tailp) / LAMBDA
        h.stats = {"MSV": (mmu, LAMBDA), "VITERBI": (vmu, LAMBDA), "FORWARD": (tau, LAMBDA)}
    return hmms


_AA = "ACDEFGHIKLMNPQRSTVWY"


def _nlp(p) -> str:
    """HMMER3 text value of a probability: -ln(p) with five decimals, '*' for 0."""
    return " ".join("*" if v <= 0.0 else f"{-math.log(v):.5f}" for v in np.asarray(p, np.float64))


def hmm_text(h: SynthHmm) -> str:
    """One model in HMMER3/f text (the format parse.rs reads): probabilities written as
    -ln(p) to five decimals, insert emissions the background, cutoffs and STATS as held."""
    out = ["HMMER3/f [3.4 | Aug 2023]", f"NAME  {h.name}"]
    if h.acc:
        out.append(f"ACC   {h.acc}")
    if h.desc:
        out.append(f"DESC  {h.desc}")
    out += [f"LENG  {h.m}", "ALPH  amino"]
    for tag, c in (("GA", h.ga), ("TC", h.tc), ("NC", h.nc)):
        if c is not None:
            out.append(f"{tag}    {c[0]:.2f} {c[1]:.2f};")
    for key in ("MSV", "VITERBI", "FORWARD"):
        if key in (h.stats or {}):
            a, b = h.stats[key]
            out.append(f"STATS LOCAL {key:<9} {a:.4f}  {b:.5f}")
    out.append("HMM          " + "        ".join(_AA))
    out.append("            m->m     m->i     m->d     i->m     i->i     d->m     d->d")
    ins = _nlp(BGN)
    out.append("  COMPO   " + _nlp(h.compo))
    out.append("          " + ins)
    out.append("          " + _nlp(h.t[0]))
    for k in range(1, h.m + 1):
        out.append(f"{k:>7}   " + _nlp(h.mat[k]) + f" {k} - - -")
        out.append("          " + ins)
        out.append("          " + _nlp(h.t[k]))
    out.append("//")
    return "\n".join(out) + "\n"


def write_hmm(path: str, hmms: List[SynthHmm]) -> None:
    """Write models as an HMMER3/f file (gzip-compressed when path ends in .gz)."""
    import gzip
    opener = gzip.open if path.endswith(".gz") else open
    with opener(path, "wt") as fh:
        for h in hmms:
            fh.write(hmm_text(h))
