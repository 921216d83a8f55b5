"""Closed-form boundary cases for the filter cascade's four decisions (pipeline.rs:164-183).

A pair's stage scores (MSV usc, null score, bias filter score fs, Viterbi vsc, Forward
fwd) are each pinned on their own.  A test's bit score is an f32, for instance
y = fl(fl(usc - null) / ln2).  Its tail probability equals the threshold F at a bit
score y* that the STATS line fixes in closed form:

  Gumbel tail (gumbel_surv, pipeline.rs:87-95): surv(y*) = F  <=>  mu = y* + ln(-ln(1 - F)) / lambda
  exponential tail (exp_logsurv, :98-104):   exp(logsurv(y*)) = F  <=>  tau = y* + ln(F) / lambda

Putting y* halfway between y and the f32 just below it makes the reference pass that
test; halfway to the f32 just above, reject it (the tail is monotone and f64 evaluation
error is far below half an f32 ulp).  The other tests of a case get STATS that put their
boundaries far away, so every case's stage follows from where its boundary was put, not
from evaluating the cascade.  Three cases sit exactly on a threshold (F1 = 1 or F2 = 0
with a tail that evaluates to exactly 1 or 0, F3 = 1 below tau), where only a strict
"reject if tail > F" passes.
"""
import math

import numpy as np

LAM = 0.693147
F = (0.02, 1e-3, 1e-5)  # distinct thresholds, so that a swap of two is visible
LOOSE = -1.0e4          # a location / tau that passes every pair
LN2F = np.float32(math.log(2.0))


def bits(score, ref):
    """The reference's bit score of a test: fl(fl(score - ref) / ln2) in f32."""
    return np.float32(np.float32(np.float32(score) - np.float32(ref)) / LN2F)


def mid_below(y):
    y = np.float32(y)
    return (float(y) + float(np.nextafter(y, np.float32(-np.inf)))) / 2.0


def mid_above(y):
    y = np.float32(y)
    return (float(y) + float(np.nextafter(y, np.float32(np.inf)))) / 2.0


def gumbel_mu(ystar, f):
    return ystar + math.log(-math.log(1.0 - f)) / LAM


def exp_tau(ystar, f):
    return ystar + math.log(f) / LAM


def ulps_apart(a, b):
    a, b = np.float32(a), np.float32(b)
    return abs(float(a) - float(b)) / float(np.spacing(max(abs(a), abs(b))))


def cases(usc, null, fs, vsc, fwd):
    """Boundary cases for one pair: (name, stats6, (f1, f2, f3), expected stage).
    Needs the MSV bit scores against the null and the bias score at least 8 ulps apart,
    so that a boundary next to one of them leaves the other clearly on its side."""
    yn, yb, yv, yf = bits(usc, null), bits(usc, fs), bits(vsc, fs), bits(fwd, fs)
    assert ulps_apart(yn, yb) >= 8
    f1, f2, f3 = F
    loose_v = (LOOSE, LAM)
    out = []

    def add(name, msv, vit, fwd_, fs_=F, expect=None):
        out.append((name, np.array([*msv, *vit, *fwd_], np.float64), fs_, expect))

    # MSV against the null score (pipeline.rs:166-168); the bias test uses the same STATS
    add("msv_null_pass", (gumbel_mu(mid_below(yn), f1), LAM), loose_v, loose_v, expect=4 if yb > yn else 1)
    add("msv_null_fail", (gumbel_mu(mid_above(yn), f1), LAM), loose_v, loose_v, expect=0)
    # MSV against the bias filter score (:170-172)
    add("msv_bias_pass", (gumbel_mu(mid_below(yb), f1), LAM), loose_v, loose_v, expect=4 if yn > yb else 0)
    add("msv_bias_fail", (gumbel_mu(mid_above(yb), f1), LAM), loose_v, loose_v, expect=1 if yn > yb else 0)
    # Viterbi against F2 (:174-177)
    add("vit_pass", loose_v, (gumbel_mu(mid_below(yv), f2), LAM), loose_v, expect=4)
    add("vit_fail", loose_v, (gumbel_mu(mid_above(yv), f2), LAM), loose_v, expect=2)
    # Forward against F3 (:179-182)
    add("fwd_pass", loose_v, loose_v, (exp_tau(mid_below(yf), f3), LAM), expect=4)
    add("fwd_fail", loose_v, loose_v, (exp_tau(mid_above(yf), f3), LAM), expect=3)
    # exactly on a threshold: a tail of exactly 1 (F1 = 1, F3 = 1) or 0 (F2 = 0) passes
    add("f1_equal", (float(yn) + 1.0e4, LAM), loose_v, loose_v, fs_=(1.0, f2, f3), expect=4)
    add("f2_equal", loose_v, (float(yv) - 1.0e4, LAM), loose_v, fs_=(f1, 0.0, f3), expect=4)
    add("f3_equal", loose_v, loose_v, (float(yf) + 1.0e4, LAM), fs_=(f1, f2, 1.0), expect=4)
    return out


def biased_seq(rng, compo, L, power=3.0):
    """Residues drawn from compo^power (renormalised): composition-biased targets whose
    bias filter score exceeds the null score."""
    p = np.asarray(compo, np.float64) ** power
    return rng.choice(20, size=L, p=p / p.sum()).astype(np.uint8)
