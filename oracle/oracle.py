"""Python access to the C oracle (sx_oracle.c) plus a plain-Python HMMER3 reader.

TEST INFRASTRUCTURE ONLY.  This module is the checker: tests/, the smoke()
check and the cpu_baseline leg of bench.py import it; the sentinel-x_amd
product never does.

The reader restates sentinel-x src/hmmer/parse.rs:107-277 (fields, cutoff
pairs, STATS lines, ``-ln p`` -> p as f32 with ``*`` = 0).
"""
from __future__ import annotations

import ctypes
import math
import os
import subprocess
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

K = 20
NCODE = 25
CODE_STAR = 24
NT = 8  # packed transition rows: MM MI MD IM II DM DD BM

_lib = None


def build(force: bool = False) -> str:
    """Compile liboracle.so with oracle/Makefile."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-B" if force else "all"], cwd=HERE, check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        f32p = ctypes.POINTER(ctypes.c_float)
        f64p = ctypes.POINTER(ctypes.c_double)
        u8p = ctypes.POINTER(ctypes.c_uint8)
        i64 = ctypes.c_int64
        L.sxo_encode.restype = ctypes.c_size_t
        L.sxo_encode.argtypes = [u8p, ctypes.c_size_t, u8p]
        L.sxo_background.restype = f32p
        L.sxo_configure.argtypes = [ctypes.c_int, f32p, f32p, f32p, f32p, f32p, f32p, f32p]
        L.sxo_xsc.argtypes = [i64, ctypes.c_int, f32p]
        L.sxo_null_score.restype = ctypes.c_float
        L.sxo_null_score.argtypes = [i64]
        L.sxo_msv.restype = ctypes.c_float
        L.sxo_msv.argtypes = [ctypes.c_int, f32p, ctypes.c_float, u8p, i64]
        L.sxo_msv16.restype = ctypes.c_int32
        L.sxo_msv16.argtypes = [ctypes.c_int, f32p, ctypes.c_float, u8p, i64]
        L.sxo_q_emit.restype = ctypes.c_int16
        L.sxo_q_emit.argtypes = [ctypes.c_float]
        L.sxo_vit16.restype = ctypes.c_int32
        L.sxo_vit16.argtypes = [ctypes.c_int, f32p, f32p, f32p, u8p, i64]
        L.sxo_q_trans.restype = ctypes.c_int16
        L.sxo_q_trans.argtypes = [ctypes.c_float]
        L.sxo_q_const.restype = ctypes.c_int32
        L.sxo_q_const.argtypes = [ctypes.c_float]
        L.sxo_viterbi.restype = ctypes.c_float
        L.sxo_viterbi.argtypes = [ctypes.c_int, f32p, f32p, f32p, u8p, i64]
        L.sxo_forward.restype = ctypes.c_double
        L.sxo_forward.argtypes = [ctypes.c_int, f32p, f32p, f32p, u8p, i64, f32p]
        L.sxo_backward.restype = ctypes.c_float
        L.sxo_backward.argtypes = [ctypes.c_int, f32p, f32p, f32p, u8p, i64, f32p]
        L.sxo_forward64.restype = ctypes.c_double
        L.sxo_forward64.argtypes = [ctypes.c_int, f32p, f32p, f32p, u8p, i64]
        L.sxo_backward64.restype = ctypes.c_double
        L.sxo_backward64.argtypes = [ctypes.c_int, f32p, f32p, f32p, u8p, i64]
        L.sxo_bias_filter.restype = ctypes.c_float
        L.sxo_bias_filter.argtypes = [ctypes.c_int, f32p, u8p, i64]
        L.sxo_gumbel_surv.restype = ctypes.c_double
        L.sxo_gumbel_surv.argtypes = [ctypes.c_double] * 3
        L.sxo_exp_logsurv.restype = ctypes.c_double
        L.sxo_exp_logsurv.argtypes = [ctypes.c_double] * 3
        L.sxo_logsum.restype = ctypes.c_float
        L.sxo_logsum.argtypes = [ctypes.c_float, ctypes.c_float]
        L.sxo_cascade.restype = ctypes.c_int
        L.sxo_cascade.argtypes = [ctypes.c_int, f32p, f32p, f32p, f32p, ctypes.c_float, f32p, f64p,
                                  ctypes.c_double, ctypes.c_double, ctypes.c_double, u8p, i64, f32p]
        L.sxo_batch_cascade.restype = ctypes.c_int
        L.sxo_batch_cascade.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double,
                                        ctypes.c_double, ctypes.c_double, i64, i64, ctypes.c_void_p,
                                        ctypes.c_void_p, ctypes.c_int]
        for name in ("sxo_batch_msv", "sxo_batch_viterbi", "sxo_batch_forward", "sxo_batch_forward64",
                     "sxo_batch_msv16", "sxo_batch_vit16", "sxo_batch_backward64"):
            fn = getattr(L, name)
            fn.restype = ctypes.c_int
            fn.argtypes = [ctypes.c_void_p, i64, i64, ctypes.c_void_p, ctypes.c_int]
        _lib = L
    return _lib


def _p(a, ct):
    return a.ctypes.data_as(ctypes.POINTER(ct))


def fp(a):
    return _p(a, ctypes.c_float)


def dp(a):
    return _p(a, ctypes.c_double)


def bp(a):
    return _p(a, ctypes.c_uint8)


# ---------------------------------------------------------------------------
# HMMER3 text reader (parse.rs restated)
# ---------------------------------------------------------------------------

@dataclass
class Hmm:
    name: str
    m: int
    mat: np.ndarray            # (M+1, 20) float32, row 0 unused
    t: np.ndarray              # (M+1, 7) float32
    acc: Optional[str] = None
    desc: Optional[str] = None
    compo: Optional[np.ndarray] = None
    ga: Optional[tuple] = None
    tc: Optional[tuple] = None
    nc: Optional[tuple] = None
    stats: dict = field(default_factory=dict)

    @property
    def calibrated(self) -> bool:
        return all(k in self.stats for k in ("MSV", "VITERBI", "FORWARD"))


def _tok_prob(tok: str) -> np.float32:
    if tok == "*":
        return np.float32(0.0)
    v = float(tok)
    if not math.isfinite(v) or v < 0.0:
        raise ValueError(f"invalid -ln(p) value {tok!r}")
    return np.float32(math.exp(-v))


def cutoff_pair(rest: str):
    toks = rest.replace(";", " ").split()
    if not toks:
        return None
    try:
        a = float(np.float32(toks[0]))
    except ValueError:
        return None
    b = a
    if len(toks) > 1:
        try:
            b = float(np.float32(toks[1]))
        except ValueError:
            b = a
    return (a, b)


def parse_hmm_text(text: str):
    lines = text.splitlines()
    pos = 0
    out = []
    n = len(lines)
    while True:
        while pos < n and not lines[pos].strip():
            pos += 1
        if pos >= n:
            break
        first = lines[pos]
        pos += 1
        if not first.startswith("HMMER3"):
            raise ValueError(f"expected 'HMMER3/x' header, found {first!r}")
        name, acc, desc, m, alph_ok = "", None, None, 0, False
        ga = tc = nc = None
        stats = {}
        while True:
            if pos >= n:
                raise ValueError("unexpected end of file (header)")
            line = lines[pos]
            pos += 1
            tr = line.lstrip()
            if tr.startswith("HMM ") or tr == "HMM" or tr.startswith("HMM\t"):
                break
            parts = tr.split(None, 1)
            tag = parts[0] if parts else ""
            rest = parts[1].strip() if len(parts) > 1 else ""
            if tag == "NAME":
                name = rest
            elif tag == "ACC":
                acc = rest
            elif tag == "DESC":
                desc = rest
            elif tag == "LENG":
                m = int(rest)
            elif tag == "ALPH":
                if rest.lower() != "amino":
                    raise ValueError(f"profile {name!r} has alphabet {rest!r}")
                alph_ok = True
            elif tag in ("GA", "TC", "NC"):
                cp = cutoff_pair(rest)
                if cp is None:
                    raise ValueError(f"bad {tag} line")
                if tag == "GA":
                    ga = cp
                elif tag == "TC":
                    tc = cp
                else:
                    nc = cp
            elif tag == "STATS":
                f = rest.split()
                if len(f) >= 4 and f[1] in ("MSV", "VITERBI", "FORWARD"):
                    stats[f[1]] = (float(f[2]), float(f[3]))
        if m == 0:
            raise ValueError(f"profile {name!r} has no LENG")
        if not alph_ok:
            raise ValueError(f"profile {name!r} has no ALPH line")
        pos += 1  # transition header
        mat = np.zeros((m + 1, K), np.float32)
        t = np.zeros((m + 1, 7), np.float32)
        compo = None
        l = lines[pos]
        pos += 1
        if l.lstrip().startswith("COMPO"):
            toks = l.split()[1:]
            if len(toks) < K:
                raise ValueError("COMPO: too few values")
            compo = np.array([_tok_prob(x) for x in toks[:K]], np.float32)
            l = lines[pos]
            pos += 1
        if len(l.split()) < K:
            raise ValueError("truncated node-0 insert line")
        toks = lines[pos].split()
        pos += 1
        if len(toks) < 7:
            raise ValueError("node 0 transitions")
        t[0] = [_tok_prob(x) for x in toks[:7]]
        for k in range(1, m + 1):
            toks = lines[pos].split()
            pos += 1
            if int(toks[0]) != k:
                raise ValueError(f"node index {toks[0]}, expected {k}")
            mat[k] = [_tok_prob(x) for x in toks[1:1 + K]]
            if len(lines[pos].split()) < K:
                raise ValueError(f"truncated insert emissions at node {k}")
            pos += 1
            toks = lines[pos].split()
            pos += 1
            if len(toks) < 7:
                raise ValueError(f"node {k} transitions")
            t[k] = [_tok_prob(x) for x in toks[:7]]
        while pos < n:
            s = lines[pos].strip()
            pos += 1
            if not s:
                continue
            if s == "//":
                break
            raise ValueError(f"expected '//', found {s!r}")
        out.append(Hmm(name=name, m=m, mat=mat, t=t, acc=acc, desc=desc, compo=compo, ga=ga, tc=tc,
                       nc=nc, stats=stats))
    return out


def parse_hmm_file(path: str):
    with open(path) as f:
        return parse_hmm_text(f.read())


# ---------------------------------------------------------------------------
# Configured profiles and DP wrappers
# ---------------------------------------------------------------------------

BG = None


def background() -> np.ndarray:
    global BG
    if BG is None:
        ptr = lib().sxo_background()
        BG = np.ctypeslib.as_array(ptr, shape=(K,)).copy()
    return BG


@dataclass
class Profile:
    name: str
    m: int
    msc: np.ndarray   # (25, M+1)
    tsc: np.ndarray   # (8, M+1)
    mo: np.ndarray
    tp: np.ndarray
    msv_tbm: float
    compo: np.ndarray
    hmm: Optional[Hmm] = None

    @property
    def stats6(self) -> np.ndarray:
        s = self.hmm.stats
        return np.array([*s["MSV"], *s["VITERBI"], *s["FORWARD"]], np.float64)


def configure(h: Hmm) -> Profile:
    m = h.m
    w = m + 1
    msc = np.empty((NCODE, w), np.float32)
    mo = np.empty((NCODE, w), np.float32)
    tsc = np.empty((NT, w), np.float32)
    tp = np.empty((NT, w), np.float32)
    tbm = np.zeros(1, np.float32)
    mat = np.ascontiguousarray(h.mat, np.float32)
    t = np.ascontiguousarray(h.t, np.float32)
    lib().sxo_configure(m, fp(mat), fp(t), fp(msc), fp(tsc), fp(mo), fp(tp), fp(tbm))
    compo = h.compo if h.compo is not None else background()
    return Profile(h.name, m, msc, tsc, mo, tp, float(tbm[0]), np.asarray(compo, np.float32), h)


def encode(seq) -> np.ndarray:
    if isinstance(seq, str):
        seq = seq.encode()
    a = np.frombuffer(seq, np.uint8).copy()
    out = np.empty(len(a), np.uint8)
    n = lib().sxo_encode(bp(a), len(a), bp(out))
    return out[:n].copy()


def xsc(l: int, multihit: bool = True) -> np.ndarray:
    x = np.empty(16, np.float32)
    lib().sxo_xsc(l, 1 if multihit else 0, fp(x))
    return x


def null_score(l: int) -> float:
    return lib().sxo_null_score(l)


def _dsq(dsq):
    return np.ascontiguousarray(dsq, np.uint8)


def msv(p: Profile, dsq) -> np.float32:
    d = _dsq(dsq)
    return np.float32(lib().sxo_msv(p.m, fp(p.msc), p.msv_tbm, bp(d), len(d)))


# Units of the quantised MSV bound (sxo_msv16): special states in 2^-22 nat.
Q_UNITS_PER_NAT = 4194304.0


def msv16(p: Profile, dsq) -> int:
    """Quantised MSV upper bound (no reference counterpart; see sx_oracle.h)."""
    d = _dsq(dsq)
    return int(lib().sxo_msv16(p.m, fp(p.msc), p.msv_tbm, bp(d), len(d)))


def vit16(p: Profile, dsq) -> int:
    """Quantised Viterbi upper bound (no reference counterpart; see sx_oracle.h)."""
    d = _dsq(dsq)
    xs = xsc(len(d))
    return int(lib().sxo_vit16(p.m, fp(p.msc), fp(p.tsc), fp(xs), bp(d), len(d)))


def viterbi(p: Profile, dsq) -> np.float32:
    d = _dsq(dsq)
    xs = xsc(len(d))
    return np.float32(lib().sxo_viterbi(p.m, fp(p.msc), fp(p.tsc), fp(xs), bp(d), len(d)))


def forward(p: Profile, dsq, specials: bool = False, multihit: bool = True, xs=None):
    """Returns the double lnf(xC*c_move)+lnscale; float32() of it is the reference score.
    xs: the 16 length-model values (default: xsc(len(dsq), multihit))."""
    d = _dsq(dsq)
    xs = xsc(len(d), multihit) if xs is None else np.ascontiguousarray(xs, np.float32)
    sp = np.empty((5, len(d) + 1), np.float32) if specials else None
    v = lib().sxo_forward(p.m, fp(p.mo), fp(p.tp), fp(xs), bp(d), len(d),
                          fp(sp) if specials else ctypes.POINTER(ctypes.c_float)())
    return (v, sp) if specials else v


def backward(p: Profile, dsq, multihit: bool = True, xs=None):
    d = _dsq(dsq)
    xs = xsc(len(d), multihit) if xs is None else np.ascontiguousarray(xs, np.float32)
    sp = np.empty((5, len(d) + 1), np.float32)
    v = lib().sxo_backward(p.m, fp(p.mo), fp(p.tp), fp(xs), bp(d), len(d), fp(sp))
    return np.float32(v), sp


def forward64(p: Profile, dsq, xs=None) -> float:
    """FP64 Forward of dsq; xs: the 16 length-model values (default: multihit, len(dsq))."""
    d = _dsq(dsq)
    xs = xsc(len(d)) if xs is None else np.ascontiguousarray(xs, np.float32)
    return lib().sxo_forward64(p.m, fp(p.mo), fp(p.tp), fp(xs), bp(d), len(d))


def backward64(p: Profile, dsq, xs=None) -> float:
    d = _dsq(dsq)
    xs = xsc(len(d)) if xs is None else np.ascontiguousarray(xs, np.float32)
    return lib().sxo_backward64(p.m, fp(p.mo), fp(p.tp), fp(xs), bp(d), len(d))


def bias_filter(p: Profile, dsq) -> np.float32:
    d = _dsq(dsq)
    return np.float32(lib().sxo_bias_filter(p.m, fp(p.compo), bp(d), len(d)))


def gumbel_surv(x, mu, lam):
    return lib().sxo_gumbel_surv(x, mu, lam)


def exp_logsurv(x, tau, lam):
    return lib().sxo_exp_logsurv(x, tau, lam)


def logsum(a, b):
    return np.float32(lib().sxo_logsum(a, b))


def cascade(p: Profile, dsq, f1=0.02, f2=1e-3, f3=1e-5):
    """pipeline.rs:159-189.  Returns (stage_reached, [msv, filtersc, vit, fwd])."""
    d = _dsq(dsq)
    out = np.empty(4, np.float32)
    st = p.stats6
    stage = lib().sxo_cascade(p.m, fp(p.msc), fp(p.tsc), fp(p.mo), fp(p.tp), p.msv_tbm, fp(p.compo),
                              dp(st), f1, f2, f3, bp(d), len(d), fp(out))
    return stage, out


# ---------------------------------------------------------------------------
# Packed many-pair drivers
# ---------------------------------------------------------------------------

class _Batch(ctypes.Structure):
    _fields_ = [("np", ctypes.c_int), ("m", ctypes.c_void_p), ("off", ctypes.c_void_p),
                ("toff", ctypes.c_void_p), ("msc", ctypes.c_void_p), ("tsc", ctypes.c_void_p),
                ("mo", ctypes.c_void_p), ("tp", ctypes.c_void_p), ("msv_tbm", ctypes.c_void_p),
                ("nt", ctypes.c_int), ("seq_off", ctypes.c_void_p), ("codes", ctypes.c_void_p)]


class Batch:
    """Packed profiles x targets for the threaded pair drivers (CPU baseline)."""

    def __init__(self, profiles, seqs):
        self.profiles = profiles
        ms = np.array([p.m for p in profiles], np.int32)
        off = np.zeros(len(profiles), np.int64)
        toff = np.zeros(len(profiles), np.int64)
        o = to = 0
        for i, p in enumerate(profiles):
            off[i] = o
            toff[i] = to
            o += NCODE * (p.m + 1)
            to += NT * (p.m + 1)
        self.m, self.off, self.toff = ms, off, toff
        self.msc = np.concatenate([p.msc.ravel() for p in profiles]).astype(np.float32)
        self.mo = np.concatenate([p.mo.ravel() for p in profiles]).astype(np.float32)
        self.tsc = np.concatenate([p.tsc.ravel() for p in profiles]).astype(np.float32)
        self.tp = np.concatenate([p.tp.ravel() for p in profiles]).astype(np.float32)
        self.tbm = np.array([p.msv_tbm for p in profiles], np.float32)
        lens = np.array([len(s) for s in seqs], np.int64)
        self.seq_off = np.zeros(len(seqs) + 1, np.int64)
        np.cumsum(lens, out=self.seq_off[1:])
        self.codes = np.concatenate([np.asarray(s, np.uint8) for s in seqs]) if seqs else np.zeros(0, np.uint8)
        self.nt = len(seqs)
        self.np_ = len(profiles)
        self._s = _Batch(self.np_, self.m.ctypes.data, self.off.ctypes.data, self.toff.ctypes.data,
                         self.msc.ctypes.data, self.tsc.ctypes.data, self.mo.ctypes.data,
                         self.tp.ctypes.data, self.tbm.ctypes.data, self.nt, self.seq_off.ctypes.data,
                         self.codes.ctypes.data)

    def _run(self, fn, dtype, begin, end, nthreads):
        if end is None:
            end = self.np_ * self.nt
        out = np.empty(end - begin, dtype)
        fn(ctypes.byref(self._s), begin, end, out.ctypes.data, nthreads)
        return out

    def msv(self, begin=0, end=None, nthreads=1):
        return self._run(lib().sxo_batch_msv, np.float32, begin, end, nthreads)

    def msv16(self, begin=0, end=None, nthreads=1):
        return self._run(lib().sxo_batch_msv16, np.int32, begin, end, nthreads)

    def vit16(self, begin=0, end=None, nthreads=1):
        return self._run(lib().sxo_batch_vit16, np.int32, begin, end, nthreads)

    def viterbi(self, begin=0, end=None, nthreads=1):
        return self._run(lib().sxo_batch_viterbi, np.float32, begin, end, nthreads)

    def forward(self, begin=0, end=None, nthreads=1):
        return self._run(lib().sxo_batch_forward, np.float64, begin, end, nthreads)

    def forward64(self, begin=0, end=None, nthreads=1):
        return self._run(lib().sxo_batch_forward64, np.float64, begin, end, nthreads)

    def backward64(self, begin=0, end=None, nthreads=1):
        return self._run(lib().sxo_batch_backward64, np.float64, begin, end, nthreads)

    def cascade(self, f1=0.02, f2=1e-3, f3=1e-5, begin=0, end=None, nthreads=1):
        """Per pair: (stage reached 0..4, [msv, filtersc, vit, fwd]) of pipeline.rs:159-189."""
        if end is None:
            end = self.np_ * self.nt
        stats = np.ascontiguousarray(np.stack([p.stats6 for p in self.profiles]), np.float64)
        compo = np.ascontiguousarray(np.stack([p.compo for p in self.profiles]), np.float32)
        stage = np.empty(end - begin, np.int32)
        sc = np.empty((end - begin, 4), np.float32)
        lib().sxo_batch_cascade(ctypes.byref(self._s), stats.ctypes.data, compo.ctypes.data, f1, f2, f3, begin, end,
                                stage.ctypes.data, sc.ctypes.data, nthreads)
        return stage, sc
