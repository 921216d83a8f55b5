"""Pins the oracle's filter-cascade decisions (sx_oracle.c sxo_cascade, the reference's
pipeline.rs:159-189) with closed-form boundary cases (tests/cascade_pins.py): each of the
four tests passes when its threshold's bit score sits half an f32 ulp below the pair's
bit score, rejects it half an ulp above, and passes when the tail equals the threshold
exactly.  Three planted mistakes -- F2 and F3 swapped, the null score where the bias
filter score belongs, and >= for > -- are each built into a scratch copy of the oracle,
and each fails these cases."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

import cascade_pins as cp
from helpers import random_hmm, random_seq

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SRC = os.path.join(ROOT, "oracle", "sx_oracle.c")


def _bind_cascade(L):
    f32p, f64p, u8p = (ctypes.POINTER(t) for t in (ctypes.c_float, ctypes.c_double, ctypes.c_uint8))
    L.sxo_cascade.restype = ctypes.c_int
    L.sxo_cascade.argtypes = [ctypes.c_int, f32p, f32p, f32p, f32p, ctypes.c_float, f32p, f64p, ctypes.c_double,
                              ctypes.c_double, ctypes.c_double, u8p, ctypes.c_int64, f32p]
    return L


def _stage(L, orc, p, d, stats6, f):
    out = np.empty(4, np.float32)
    d = np.ascontiguousarray(d, np.uint8)
    return L.sxo_cascade(p.m, orc.fp(p.msc), orc.fp(p.tsc), orc.fp(p.mo), orc.fp(p.tp), p.msv_tbm, orc.fp(p.compo),
                         orc.dp(np.ascontiguousarray(stats6)), f[0], f[1], f[2], orc.bp(d), len(d), orc.fp(out))


@pytest.fixture(scope="module")
def pin_pairs(orc):
    """(profile, target, stage scores) with the two MSV bit scores at least 8 ulps apart:
    background targets (bias score below the null score) and composition-biased ones
    (above it)."""
    rng = np.random.default_rng(77)
    out = []
    for seed, m in ((31, 40), (32, 120)):
        p = orc.configure(random_hmm(orc, m, seed))
        for i in range(12):
            L = int(rng.integers(30, 400))
            d = cp.biased_seq(rng, p.compo, L) if i % 2 else random_seq(rng, L, degen_frac=0.02)
            sc = (orc.msv(p, d), np.float32(orc.null_score(len(d))), orc.bias_filter(p, d), orc.viterbi(p, d),
                  np.float32(orc.forward(p, d)))
            if cp.ulps_apart(cp.bits(sc[0], sc[1]), cp.bits(sc[0], sc[2])) >= 8:
                out.append((p, d, sc))
    yn_gt_yb = [cp.bits(s[0], s[1]) > cp.bits(s[0], s[2]) for _, _, s in out]
    assert any(yn_gt_yb) and not all(yn_gt_yb), "need both orders of the two MSV bit scores"
    return out


def _failures(L, orc, pairs):
    bad = []
    for k, (p, d, sc) in enumerate(pairs):
        for name, stats6, f, want in cp.cases(*sc):
            got = _stage(L, orc, p, d, stats6, f)
            if got != want:
                bad.append((k, name, want, got))
    return bad


def test_oracle_cascade_meets_every_boundary_case(orc, pin_pairs):
    L = _bind_cascade(orc.lib())
    assert _failures(L, orc, pin_pairs) == []


def test_boundary_cases_see_the_tails_they_place(orc, pin_pairs):
    """The placed boundaries straddle each bit score as the closed forms say (checked with
    the oracle's tail functions, themselves pinned by pipeline.rs:600-604)."""
    for _, _, sc in pin_pairs[:4]:
        yn = cp.bits(sc[0], sc[1])
        mu = cp.gumbel_mu(cp.mid_below(yn), cp.F[0])
        assert orc.gumbel_surv(float(yn), mu, cp.LAM) < cp.F[0]
        assert orc.gumbel_surv(float(np.nextafter(yn, np.float32(-np.inf))), mu, cp.LAM) > cp.F[0]
        yf = cp.bits(sc[4], sc[2])
        tau = cp.exp_tau(cp.mid_below(yf), cp.F[2])
        assert np.exp(orc.exp_logsurv(float(yf), tau, cp.LAM)) < cp.F[2]
        assert np.exp(orc.exp_logsurv(float(np.nextafter(yf, np.float32(-np.inf))), tau, cp.LAM)) > cp.F[2]


@pytest.mark.parametrize("plant,what", [(1, "F2 and F3 swapped"), (2, "null score for the bias score"),
                                        (3, ">= for >")])
def test_boundary_cases_catch_a_planted_mistake(orc, pin_pairs, tmp_path, plant, what):
    so = str(tmp_path / f"liboracle_plant{plant}.so")
    subprocess.run(["gcc", "-O2", "-mavx2", "-ffp-contract=off", "-fno-fast-math", "-fPIC", "-std=gnu11", "-shared",
                    f"-DSXO_PLANT={plant}", "-o", so, ORACLE_SRC, "-lm", "-lpthread"], check=True)
    L = _bind_cascade(ctypes.CDLL(so))
    bad = _failures(L, orc, pin_pairs)
    assert bad, f"the planted mistake ({what}) passed every boundary case"
