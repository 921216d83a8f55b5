"""sentinel-x_amd: MI355X implementation of sentinel-x's src/hmmer Plan7 search path.

The compute lives in libsxhmmer.so (HIP kernels for gfx950 behind the C ABI of
include/sx_hmmer.h).  This module is a thin ctypes binding used by the tests, the
benchmark and Python callers; it never falls back to a CPU implementation: if the
library or a HIP device is missing, it raises.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from typing import Optional, Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsxhmmer.so")

STAGE_MSV, STAGE_VITERBI, STAGE_FORWARD, STAGE_BACKWARD = 0, 1, 2, 3
# int16 MSV upper bound (no reference counterpart): int32 scores in units of 2^-22 nat,
# at least 2^22 times the f32 MSV score (see sx_hmmer.h for a bound whose cells saturated)
STAGE_MSV16 = 4
# int16 Viterbi upper bound (no reference counterpart), in the same units
STAGE_VIT16 = 5
MSV16_UNITS_PER_NAT = 4194304.0
FORM_PLAIN, FORM_BATCHED = 0, 1
STAGE_NAMES = {0: "msv", 1: "viterbi", 2: "forward", 3: "backward", 4: "msv16", 5: "vit16"}


def _stage_dtype(stage: int):
    return {STAGE_MSV16: np.int32, STAGE_VIT16: np.int32, STAGE_FORWARD: np.float64,
            STAGE_BACKWARD: np.float64}.get(stage, np.float32)

# Every entry point declared in include/sx_hmmer.h.
EXPORTS = [
    "sx_create", "sx_destroy", "sx_last_error", "sx_set_targets", "sx_set_profiles", "sx_load_hmm_file",
    "sx_num_profiles", "sx_num_targets", "sx_profile_tables", "sx_score_all", "sx_score_pairs", "sx_specials",
    "sx_run_cascade", "sx_cascade_records", "sx_cascade_launches", "sx_autotune", "sx_shape_for", "sx_set_shapes", "sx_get_shapes",
    "sx_autotune_log", "sx_search", "sx_search_host", "sx_results_size", "sx_results_stats", "sx_results_copy",
    "sx_results_free", "sx_encode", "sx_profile_info", "sx_score_pairs_noalloc", "sx_search_hmm_file",
    "sx_results_models", "sx_results_model", "sx_write_profile_cache", "sx_convert_hmm_file", "sx_rescore",
    "sx_rescore_host",
]
NOALLOC_MAX_PAIRS = 1 << 18  # SX_NOALLOC_MAX_PAIRS


class SxError(RuntimeError):
    pass


def build(force: bool = True, jobs: int = 4) -> str:
    """Compile libsxhmmer.so for gfx950 with the package Makefile."""
    cmd = ["make", f"-j{jobs}"] + (["-B"] if force else [])
    subprocess.run(cmd, cwd=HERE, check=True)
    return LIB_PATH


class sx_hmm(ctypes.Structure):
    _fields_ = [
        ("name", ctypes.c_char_p), ("acc", ctypes.c_char_p), ("desc", ctypes.c_char_p), ("m", ctypes.c_int32),
        ("mat", ctypes.POINTER(ctypes.c_float)), ("t", ctypes.POINTER(ctypes.c_float)),
        ("compo", ctypes.POINTER(ctypes.c_float)),
        ("has_ga", ctypes.c_int32), ("has_tc", ctypes.c_int32), ("has_nc", ctypes.c_int32),
        ("ga", ctypes.c_float * 2), ("tc", ctypes.c_float * 2), ("nc", ctypes.c_float * 2),
        ("has_stats", ctypes.c_int32), ("stats", ctypes.c_double * 6),
    ]


class sx_pair_scores(ctypes.Structure):
    _fields_ = [("profile", ctypes.c_uint32), ("target", ctypes.c_uint32), ("msv", ctypes.c_float),
                ("bias", ctypes.c_float), ("vit", ctypes.c_float), ("fwd", ctypes.c_float),
                ("fwd64", ctypes.c_double), ("stage", ctypes.c_int32), ("flags", ctypes.c_int32)]


# sx_pair_scores.flags: the bias score was recomputed on the host (the device's error
# bound could not certify its f32 rounding)
REC_BIAS_HOST = 1


class sx_cascade_stats(ctypes.Structure):
    _fields_ = [("f1", ctypes.c_double), ("f2", ctypes.c_double), ("f3", ctypes.c_double),
                ("use_graph", ctypes.c_int32), ("bounds", ctypes.c_int32), ("n_pairs", ctypes.c_uint64),
                ("n_msv_pass", ctypes.c_uint64), ("n_bias_pass", ctypes.c_uint64),
                ("n_vit_pass", ctypes.c_uint64), ("n_fwd_pass", ctypes.c_uint64),
                ("ms_msv", ctypes.c_double), ("ms_bias", ctypes.c_double), ("ms_vit", ctypes.c_double),
                ("ms_fwd", ctypes.c_double), ("ms_total", ctypes.c_double),
                ("cells_msv", ctypes.c_uint64), ("cells_vit", ctypes.c_uint64), ("cells_fwd", ctypes.c_uint64),
                ("n_msv_cand", ctypes.c_uint64), ("ms_msv_bound", ctypes.c_double),
                ("cells_msv_exact", ctypes.c_uint64), ("n_vit_cand", ctypes.c_uint64),
                ("ms_vit_bound", ctypes.c_double), ("n_bias_recheck", ctypes.c_uint64),
                ("ms_recheck", ctypes.c_double)]


class sx_search_cfg(ctypes.Structure):
    _fields_ = [("cutoff", ctypes.c_int32), ("filters", ctypes.c_int32), ("null2", ctypes.c_int32),
                ("threads", ctypes.c_int32), ("seq_e", ctypes.c_double), ("dom_e", ctypes.c_double),
                ("seq_t", ctypes.c_float), ("dom_t", ctypes.c_float), ("f1", ctypes.c_double),
                ("f2", ctypes.c_double), ("f3", ctypes.c_double), ("z", ctypes.c_double),
                ("host_rescore", ctypes.c_int32), ("pad", ctypes.c_int32)]


ENVELOPE_DTYPE = np.dtype([("profile", "<u4"), ("target", "<u4"), ("i", "<i4"), ("j", "<i4"), ("null2", "<i4")])
ENVELOPE_SCORE_DTYPE = np.dtype([("envsc", "<f4"), ("bcksc", "<f4"), ("status", "<i4"), ("oa", "<i4"),
                                 ("iali", "<i4"), ("jali", "<i4"), ("kali", "<i4"), ("mali", "<i4"),
                                 ("null2", "<f4", (20,))])

HIT_DTYPE = np.dtype([("profile", "<u4"), ("target", "<u4"), ("qlen", "<u4"), ("tlen", "<u4"), ("score", "<f4"),
                      ("bias", "<f4"), ("evalue", "<f8"), ("nexpected", "<f4"), ("ndom", "<u4"),
                      ("dom_first", "<u8")])
DOMAIN_DTYPE = np.dtype([("score", "<f4"), ("bias", "<f4"), ("c_evalue", "<f8"), ("i_evalue", "<f8"),
                         ("hmm_from", "<u4"), ("hmm_to", "<u4"), ("ali_from", "<u4"), ("ali_to", "<u4"),
                         ("env_from", "<u4"), ("env_to", "<u4"), ("reported", "<i4"), ("pad", "<i4")])

PAIR_DTYPE = np.dtype([("profile", "<u4"), ("target", "<u4"), ("msv", "<f4"), ("bias", "<f4"), ("vit", "<f4"),
                       ("fwd", "<f4"), ("fwd64", "<f8"), ("stage", "<i4"), ("flags", "<i4")])

LAUNCH_DTYPE = np.dtype([("stage", "<i4"), ("q", "<i4"), ("g", "<i4"), ("hoist", "<i4"), ("ms", "<f8"),
                         ("cells", "<u8")])

_lib = None


def lib():
    """Load libsxhmmer.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise SxError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() or `make` in {HERE}")
        L = ctypes.CDLL(LIB_PATH)
        vp, u32, u64, i32 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int32
        L.sx_create.argtypes = [ctypes.c_int, ctypes.POINTER(vp)]
        L.sx_destroy.argtypes = [vp]
        L.sx_destroy.restype = None
        L.sx_last_error.argtypes = [vp]
        L.sx_last_error.restype = ctypes.c_char_p
        L.sx_set_targets.argtypes = [vp, vp, vp, u32]
        L.sx_set_profiles.argtypes = [vp, ctypes.POINTER(sx_hmm), u32]
        L.sx_load_hmm_file.argtypes = [vp, ctypes.c_char_p, ctypes.POINTER(u32)]
        L.sx_num_profiles.argtypes = [vp]
        L.sx_num_profiles.restype = u32
        L.sx_num_targets.argtypes = [vp]
        L.sx_num_targets.restype = u32
        L.sx_profile_tables.argtypes = [vp, u32, ctypes.POINTER(i32), vp, vp, vp, vp, vp]
        L.sx_encode.argtypes = [vp, ctypes.c_size_t, vp, ctypes.POINTER(ctypes.c_size_t)]
        cpp = ctypes.POINTER(ctypes.c_char_p)
        L.sx_profile_info.argtypes = [vp, u32, cpp, cpp, cpp, ctypes.POINTER(i32)]
        L.sx_score_all.argtypes = [vp, ctypes.c_int, ctypes.c_int, vp, ctypes.POINTER(ctypes.c_double)]
        L.sx_score_pairs.argtypes = [vp, ctypes.c_int, ctypes.c_int, vp, vp, u64, vp, ctypes.POINTER(ctypes.c_double)]
        L.sx_score_pairs_noalloc.argtypes = [vp, ctypes.c_int, ctypes.c_int, vp, vp, u64, vp,
                                             ctypes.POINTER(ctypes.c_double)]
        L.sx_specials.argtypes = [vp, ctypes.c_int, vp, vp, u64, vp, vp]
        L.sx_rescore.argtypes = [vp, vp, u64, vp]
        L.sx_rescore_host.argtypes = [vp, vp, u64, vp]
        L.sx_run_cascade.argtypes = [vp, ctypes.c_int, ctypes.POINTER(sx_cascade_stats), vp, u64,
                                     ctypes.POINTER(u64)]
        L.sx_cascade_records.argtypes = [vp, vp, u64, ctypes.POINTER(u64)]
        L.sx_cascade_launches.argtypes = [vp, u32, ctypes.POINTER(u32), vp]
        L.sx_autotune.argtypes = [vp, ctypes.c_int, ctypes.c_double]
        L.sx_search.argtypes = [vp, ctypes.POINTER(sx_search_cfg), ctypes.POINTER(vp)]
        L.sx_search_host.argtypes = [ctypes.POINTER(sx_hmm), u32, vp, vp, u32, ctypes.POINTER(sx_search_cfg), u64,
                                     vp, vp, vp, vp, vp, vp, ctypes.POINTER(vp)]
        L.sx_results_size.argtypes = [vp, ctypes.POINTER(u64), ctypes.POINTER(u64)]
        L.sx_results_copy.argtypes = [vp, vp, vp]
        L.sx_results_stats.argtypes = [vp, ctypes.POINTER(SearchStats)]
        L.sx_results_free.argtypes = [vp]
        L.sx_results_free.restype = None
        L.sx_search_hmm_file.argtypes = [vp, ctypes.c_char_p, ctypes.POINTER(sx_search_cfg), u32, ctypes.POINTER(vp)]
        L.sx_results_models.argtypes = [vp, ctypes.POINTER(u32)]
        L.sx_results_model.argtypes = [vp, u32, ctypes.POINTER(ctypes.c_char_p), ctypes.POINTER(ctypes.c_char_p),
                                       ctypes.POINTER(i32)]
        L.sx_write_profile_cache.argtypes = [vp, ctypes.c_char_p]
        L.sx_convert_hmm_file.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(u32)]
        L.sx_shape_for.argtypes = [vp, ctypes.c_int, ctypes.c_int, i32, ctypes.POINTER(i32), ctypes.POINTER(i32),
                                   ctypes.POINTER(i32)]
        L.sx_set_shapes.argtypes = [vp, ctypes.c_int, u32, vp, vp, vp, vp]
        L.sx_get_shapes.argtypes = [vp, ctypes.c_int, u32, ctypes.POINTER(u32), vp, vp, vp, vp]
        L.sx_autotune_log.argtypes = [vp, ctypes.c_int, u32, ctypes.POINTER(u32), vp, vp, vp, vp, vp]
        for name in EXPORTS:
            getattr(L, name)  # raises AttributeError if the library lacks a declared entry point
        _lib = L
    return _lib


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else ctypes.c_void_p(0)


def _hmm_array(hmms):
    """An sx_hmm array viewing the models' arrays (and the list of objects that must stay alive)."""
    arr = (sx_hmm * max(1, len(hmms)))()
    keep = [arr]
    for i, h in enumerate(hmms):
        mat = np.ascontiguousarray(h.mat, np.float32)
        t = np.ascontiguousarray(h.t, np.float32)
        keep += [mat, t]
        e = arr[i]
        nm = h.name.encode()
        keep.append(nm)
        e.name = nm
        e.acc = getattr(h, "acc", None).encode() if getattr(h, "acc", None) else None
        e.desc = getattr(h, "desc", None).encode() if getattr(h, "desc", None) else None
        e.m = h.m
        e.mat = mat.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
        e.t = t.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
        if getattr(h, "compo", None) is not None:
            c = np.ascontiguousarray(h.compo, np.float32)
            keep.append(c)
            e.compo = c.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
        for nm2 in ("ga", "tc", "nc"):
            v = getattr(h, nm2, None)
            if v is not None:
                setattr(e, "has_" + nm2, 1)
                getattr(e, nm2)[0], getattr(e, nm2)[1] = v
        st = getattr(h, "stats", {}) or {}
        bits = 0
        for b, key in enumerate(("MSV", "VITERBI", "FORWARD")):
            if key in st:
                bits |= 1 << b
                e.stats[2 * b], e.stats[2 * b + 1] = st[key]
        e.has_stats = bits
    return arr, keep


CUTOFFS = {"evalue": 0, "bits": 1, "ga": 2, "nc": 3, "tc": 4}


def search_config(cutoff="evalue", seq_e=10.0, dom_e=10.0, seq_t=0.0, dom_t=0.0, filters=True, null2=True,
                  f1=0.02, f2=1e-3, f3=1e-5, z=None, threads=0, host_rescore=False):
    """SearchConfig (search.rs:42-47) with PipelineConfig (pipeline.rs:27-42) as an sx_search_cfg.
    host_rescore: envelope rescoring on host threads instead of the device."""
    c = sx_search_cfg()
    c.host_rescore = int(bool(host_rescore))
    c.cutoff = CUTOFFS[cutoff]
    c.filters, c.null2, c.threads = int(bool(filters)), int(bool(null2)), int(threads)
    c.seq_e, c.dom_e, c.seq_t, c.dom_t = seq_e, dom_e, seq_t, dom_t
    c.f1, c.f2, c.f3 = f1, f2, f3
    c.z = float(z) if z else 0.0
    return c


class SearchStats(ctypes.Structure):
    """sx_search_stats: where a search spent its time (ms) and how much reached the host."""
    _fields_ = [("n_pairs", ctypes.c_uint64), ("n_rows", ctypes.c_uint64), ("ms_filters", ctypes.c_double),
                ("ms_specials", ctypes.c_double), ("ms_tail", ctypes.c_double), ("ms_report", ctypes.c_double),
                ("n_regions", ctypes.c_uint64), ("bytes_to_host", ctypes.c_uint64), ("n_models", ctypes.c_uint64),
                ("ms_ingest", ctypes.c_double), ("ms_ingest_wait", ctypes.c_double),
                ("ms_rescore", ctypes.c_double), ("n_envelopes", ctypes.c_uint64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def _take_results(L, r, stats=None, models=None):
    """(hits, domains) of an sx_results, which it frees; the stats and the model list
    ((name, acc, m) per profile index) go into `stats` / `models` when given."""
    nh, nd = ctypes.c_uint64(), ctypes.c_uint64()
    try:
        L.sx_results_size(r, ctypes.byref(nh), ctypes.byref(nd))
        hits = np.zeros(nh.value, HIT_DTYPE)
        doms = np.zeros(nd.value, DOMAIN_DTYPE)
        L.sx_results_copy(r, _ptr(hits), _ptr(doms))
        if stats is not None:
            L.sx_results_stats(r, ctypes.byref(stats))
        if models is not None:
            n = ctypes.c_uint32()
            L.sx_results_models(r, ctypes.byref(n))
            nm, ac, m = ctypes.c_char_p(), ctypes.c_char_p(), ctypes.c_int32()
            for i in range(n.value):
                L.sx_results_model(r, i, ctypes.byref(nm), ctypes.byref(ac), ctypes.byref(m))
                models.append((nm.value.decode(), ac.value.decode() if ac.value is not None else None, m.value))
    finally:
        L.sx_results_free(r)
    return hits, doms


def convert_hmm_file(hmm_path: str, cache_path: str, threads: int = 0) -> int:
    """sx_convert_hmm_file: parse and configure every model of an HMMER3 file (text or
    gzip) and write the packed profile cache; returns the number of models.  Needs no
    device."""
    n = ctypes.c_uint32()
    L = lib()
    rc = L.sx_convert_hmm_file(hmm_path.encode(), cache_path.encode(), int(threads), ctypes.byref(n))
    if rc != 0:
        raise SxError(f"sx_convert_hmm_file failed ({rc}): {L.sx_last_error(None).decode()}")
    return n.value


def encode(seq) -> np.ndarray:
    """alphabet::encode (alphabet.rs:91-93) through sx_encode: ASCII residues (bytes or
    str) to codes 0..24, whitespace dropped.  Needs no device."""
    if isinstance(seq, str):
        seq = seq.encode()
    a = np.frombuffer(bytes(seq), np.uint8)
    out = np.empty(max(1, len(a)), np.uint8)
    n = ctypes.c_size_t()
    rc = lib().sx_encode(_ptr(a) if len(a) else None, len(a), _ptr(out), ctypes.byref(n))
    if rc != 0:
        raise SxError(f"sx_encode failed ({rc})")
    return out[:n.value].copy()


def search_host(hmms, seqs, prof, tgt, fwd, bwd, fsp, bsp, **cfg):
    """The host part of the search alone (sx_search_host): pairs (prof[i], tgt[i]) with their
    whole-sequence Forward / Backward scores and special rows (fsp / bsp: for each pair a
    (5, L+1) array, rows N B E J C).  Needs no device."""
    L = lib()
    arr, keep = _hmm_array(hmms)
    lens = np.array([len(x) for x in seqs], np.uint64)
    off = np.zeros(len(seqs) + 1, np.uint64)
    np.cumsum(lens, out=off[1:])
    codes = np.concatenate([np.asarray(x, np.uint8) for x in seqs]) if len(seqs) else np.zeros(1, np.uint8)
    if codes.size == 0:
        codes = np.zeros(1, np.uint8)
    prof = np.ascontiguousarray(prof, np.uint32)
    tgt = np.ascontiguousarray(tgt, np.uint32)
    fwd = np.ascontiguousarray(fwd, np.float64)
    bwd = np.ascontiguousarray(bwd, np.float64)
    f = np.ascontiguousarray(np.concatenate([np.asarray(x, np.float32).ravel() for x in fsp]) if len(fsp)
                             else np.zeros(1, np.float32))
    b = np.ascontiguousarray(np.concatenate([np.asarray(x, np.float32).ravel() for x in bsp]) if len(bsp)
                             else np.zeros(1, np.float32))
    c = search_config(**cfg)
    r = ctypes.c_void_p()
    rc = L.sx_search_host(arr, len(hmms), _ptr(codes), _ptr(off), len(seqs), ctypes.byref(c), len(prof), _ptr(prof),
                          _ptr(tgt), _ptr(fwd), _ptr(bwd), _ptr(f), _ptr(b), ctypes.byref(r))
    if rc != 0:
        raise SxError(f"error {rc}: {L.sx_last_error(None).decode()}")
    return _take_results(L, r)



class Context:
    """One HIP device with resident profiles and targets (see include/sx_hmmer.h)."""

    def __init__(self, device: int = 0):
        self._L = lib()
        h = ctypes.c_void_p()
        rc = self._L.sx_create(device, ctypes.byref(h))
        if rc != 0:
            raise SxError(f"sx_create failed ({rc}): {self._L.sx_last_error(None).decode()}")
        self._h = h
        self._keep = []
        self.lengths = np.zeros(0, np.int64)
        self.n_profiles = 0

    def close(self):
        if getattr(self, "_h", None):
            self._L.sx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _chk(self, rc):
        if rc != 0:
            raise SxError(f"error {rc}: {self._L.sx_last_error(self._h).decode()}")

    # -- data --
    def set_targets(self, seqs: Sequence[np.ndarray]):
        lens = np.array([len(s) for s in seqs], np.uint64)
        off = np.zeros(len(seqs) + 1, np.uint64)
        np.cumsum(lens, out=off[1:])
        codes = np.concatenate([np.asarray(s, np.uint8) for s in seqs]) if len(seqs) else np.zeros(1, np.uint8)
        if codes.size == 0:
            codes = np.zeros(1, np.uint8)
        self._chk(self._L.sx_set_targets(self._h, _ptr(codes), _ptr(off), len(seqs)))
        self.lengths = lens.astype(np.int64)

    def set_profiles(self, hmms):
        """hmms: objects with name, m, mat (M+1,20), t (M+1,7), compo, ga/tc/nc, stats dict."""
        arr, keep = _hmm_array(hmms)
        self._chk(self._L.sx_set_profiles(self._h, arr, len(hmms)))
        self.n_profiles = len(hmms)

    def search(self, **cfg):
        """hmmsearch-equivalent search of the loaded profiles against the loaded targets
        (see sx_search).  Keywords as search_config().  Returns (hits, domains) structured
        arrays; hit h's domains are domains[h['dom_first'] : h['dom_first'] + h['ndom']].
        The call's phase times (sx_search_stats) are left in self.last_search_stats."""
        c = search_config(**cfg)
        r = ctypes.c_void_p()
        self._chk(self._L.sx_search(self._h, ctypes.byref(c), ctypes.byref(r)))
        st = SearchStats()
        out = _take_results(self._L, r, st)
        self.last_search_stats = st.as_dict()
        return out

    def load_hmm_file(self, path: str) -> int:
        """Append the models of an HMMER3 file (text, gzip or a packed profile cache)."""
        n = ctypes.c_uint32()
        self._chk(self._L.sx_load_hmm_file(self._h, path.encode(), ctypes.byref(n)))
        self.n_profiles = self._L.sx_num_profiles(self._h)
        return n.value

    def write_profile_cache(self, path: str):
        """sx_write_profile_cache: the loaded profiles as a packed profile cache."""
        self._chk(self._L.sx_write_profile_cache(self._h, path.encode()))

    def search_hmm_file(self, path: str, batch: int = 0, **cfg):
        """search_hmm_file_models (search.rs:252-271) through sx_search_hmm_file: every model
        of the file against the loaded targets, `batch` at a time (0: 2048; the reference
        takes 256, and the results do not depend on it).  Returns (hits,
        domains, models) with models[i] = (name, acc, m) of profile index i; the phase
        times go to self.last_search_stats."""
        c = search_config(**cfg)
        r = ctypes.c_void_p()
        self._chk(self._L.sx_search_hmm_file(self._h, path.encode(), ctypes.byref(c), int(batch), ctypes.byref(r)))
        st, models = SearchStats(), []
        hits, doms = _take_results(self._L, r, st, models)
        self.last_search_stats = st.as_dict()
        self.n_profiles = self._L.sx_num_profiles(self._h)
        return hits, doms, models

    def profile_info(self, p: int):
        """ModelInfo (search.rs:245-249) of profile p: (name, acc, desc, m); acc and desc
        are None when the model has none."""
        nm, ac, de = ctypes.c_char_p(), ctypes.c_char_p(), ctypes.c_char_p()
        m = ctypes.c_int32()
        self._chk(self._L.sx_profile_info(self._h, p, ctypes.byref(nm), ctypes.byref(ac), ctypes.byref(de),
                                          ctypes.byref(m)))
        dec = lambda v: v.value.decode() if v.value is not None else None
        return dec(nm), dec(ac), dec(de), m.value

    def profile_tables(self, p: int):
        m = ctypes.c_int32()
        self._chk(self._L.sx_profile_tables(self._h, p, ctypes.byref(m), None, None, None, None, None))
        w = m.value + 1
        msc = np.empty((25, w), np.float32)
        mo = np.empty((25, w), np.float32)
        tsc = np.empty((8, w), np.float32)
        tp = np.empty((8, w), np.float32)
        tbm = np.zeros(1, np.float32)
        self._chk(self._L.sx_profile_tables(self._h, p, ctypes.byref(m), _ptr(msc), _ptr(tsc), _ptr(mo), _ptr(tp),
                                            _ptr(tbm)))
        return msc, tsc, mo, tp, float(tbm[0])

    # -- compute --
    def score_all(self, stage: int, form: int = FORM_BATCHED, want: bool = True):
        nt = len(self.lengths)
        dt = _stage_dtype(stage)
        out = np.empty((self.n_profiles, nt), dt) if want else None
        ms = ctypes.c_double()
        self._chk(self._L.sx_score_all(self._h, stage, form, _ptr(out), ctypes.byref(ms)))
        return out, ms.value

    def score_pairs(self, stage: int, prof, tgt, form: int = FORM_BATCHED):
        prof = np.ascontiguousarray(prof, np.uint32)
        tgt = np.ascontiguousarray(tgt, np.uint32)
        dt = _stage_dtype(stage)
        out = np.empty(len(prof), dt)
        ms = ctypes.c_double()
        self._chk(self._L.sx_score_pairs(self._h, stage, form, _ptr(prof), _ptr(tgt), len(prof), _ptr(out),
                                         ctypes.byref(ms)))
        return out, ms.value

    def score_pairs_noalloc(self, stage: int, prof, tgt, form: int = FORM_BATCHED, out=None):
        """score_pairs through sx_score_pairs_noalloc (the library allocates nothing during
        the call; at most NOALLOC_MAX_PAIRS pairs).  out, if given, is the result array."""
        prof = np.ascontiguousarray(prof, np.uint32)
        tgt = np.ascontiguousarray(tgt, np.uint32)
        if out is None:
            out = np.empty(len(prof), _stage_dtype(stage))
        ms = ctypes.c_double()
        self._chk(self._L.sx_score_pairs_noalloc(self._h, stage, form, _ptr(prof), _ptr(tgt), len(prof), _ptr(out),
                                                 ctypes.byref(ms)))
        return out, ms.value

    def specials(self, stage: int, prof, tgt):
        """Returns (list of (5, L+1) special-row arrays, scores)."""
        prof = np.ascontiguousarray(prof, np.uint32)
        tgt = np.ascontiguousarray(tgt, np.uint32)
        L = self.lengths[tgt]
        rows = int((L + 1).sum())
        sp = np.empty(rows * 5, np.float32)
        sc = np.empty(len(prof), np.float64)
        self._chk(self._L.sx_specials(self._h, stage, _ptr(prof), _ptr(tgt), len(prof), _ptr(sp), _ptr(sc)))
        out, o = [], 0
        for l in L:
            out.append(sp[o * 5:(o + l + 1) * 5].reshape(5, l + 1))
            o += l + 1
        return out, sc

    def rescore(self, envelopes, on_host: bool = False):
        """Envelope rescoring (sx_rescore, or sx_rescore_host with on_host) of an
        ENVELOPE_DTYPE array (or (profile, target, i, j, null2) tuples); returns an
        ENVELOPE_SCORE_DTYPE array."""
        env = np.ascontiguousarray(np.asarray(envelopes, ENVELOPE_DTYPE) if not isinstance(envelopes, np.ndarray)
                                   else envelopes.astype(ENVELOPE_DTYPE))
        out = np.zeros(len(env), ENVELOPE_SCORE_DTYPE)
        fn = self._L.sx_rescore_host if on_host else self._L.sx_rescore
        self._chk(fn(self._h, _ptr(env), len(env), _ptr(out)))
        return out

    def run_cascade(self, form: int = FORM_BATCHED, f1=0.02, f2=1e-3, f3=1e-5, use_graph: bool = True,
                    cap: Optional[int] = None, msv_bound: bool = True, vit_bound: bool = False):
        """Returns (stats, records) for every MSV survivor (see sx_run_cascade).  With cap the
        records are truncated to cap; without it all of them are fetched after the run.
        msv_bound / vit_bound: run the int16 MSV / Viterbi bound first and the f32 stage only
        on its candidates (records the Viterbi bound rejects carry vit = NaN).  The Viterbi
        bound is off by default: on gfx950 it is no faster than the f32 Viterbi."""
        st = sx_cascade_stats()
        st.f1, st.f2, st.f3 = f1, f2, f3
        st.use_graph = 1 if use_graph else 0
        st.bounds = (1 if msv_bound else 0) | (2 if vit_bound else 0)
        n = ctypes.c_uint64()
        if cap is not None:
            out = np.empty(cap, PAIR_DTYPE)
            self._chk(self._L.sx_run_cascade(self._h, form, ctypes.byref(st), _ptr(out), cap, ctypes.byref(n)))
            return st, out[:min(cap, n.value)].copy()
        self._chk(self._L.sx_run_cascade(self._h, form, ctypes.byref(st), None, 0, ctypes.byref(n)))
        return st, self.cascade_records(int(n.value))

    def cascade_records(self, cap: Optional[int] = None):
        """Records of the last run_cascade (all of them, or the first cap)."""
        n = ctypes.c_uint64()
        if cap is None:
            self._chk(self._L.sx_cascade_records(self._h, None, 0, ctypes.byref(n)))
            cap = int(n.value)
        out = np.empty(cap, PAIR_DTYPE)
        self._chk(self._L.sx_cascade_records(self._h, _ptr(out), cap, ctypes.byref(n)))
        return out[:min(cap, n.value)]

    def cascade_launches(self):
        """Per-launch device time of the last eager batched run_cascade (see sx_cascade_launches)."""
        n = ctypes.c_uint32()
        self._chk(self._L.sx_cascade_launches(self._h, 0, ctypes.byref(n), None))
        out = np.zeros(n.value, LAUNCH_DTYPE)
        if n.value:
            self._chk(self._L.sx_cascade_launches(self._h, n.value, ctypes.byref(n), _ptr(out)))
        return out

    def autotune(self, stage: int, budget_ms: float = 2000.0):
        self._chk(self._L.sx_autotune(self._h, stage, budget_ms))

    def set_shapes(self, stage: int, table):
        """table: rows (max_m, q, g[, hoist]) in increasing max_m (the batched form's shape table)."""
        mm = np.array([r[0] for r in table], np.int32)
        q = np.array([r[1] for r in table], np.int32)
        g = np.array([r[2] for r in table], np.int32)
        h = np.array([r[3] if len(r) > 3 else 0 for r in table], np.int32)
        self._chk(self._L.sx_set_shapes(self._h, stage, len(table), _ptr(mm), _ptr(q), _ptr(g), _ptr(h)))

    def get_shapes(self, stage: int):
        n = ctypes.c_uint32()
        cap = 64
        mm, q, g, h = (np.zeros(cap, np.int32) for _ in range(4))
        self._chk(self._L.sx_get_shapes(self._h, stage, cap, ctypes.byref(n), _ptr(mm), _ptr(q), _ptr(g), _ptr(h)))
        k = min(n.value, cap)
        return [(int(mm[i]), int(q[i]), int(g[i]), int(h[i])) for i in range(k)]

    def autotune_log(self, stage: int):
        """Rows (bucket max_m, q, g, hoist, gcups) measured by the last autotune of `stage`."""
        n = ctypes.c_uint32()
        self._chk(self._L.sx_autotune_log(self._h, stage, 0, ctypes.byref(n), None, None, None, None, None))
        k = n.value
        mm, q, g, h = (np.zeros(max(k, 1), np.int32) for _ in range(4))
        gc = np.zeros(max(k, 1), np.float64)
        self._chk(self._L.sx_autotune_log(self._h, stage, k, ctypes.byref(n), _ptr(mm), _ptr(q), _ptr(g), _ptr(h),
                                          _ptr(gc)))
        return [(int(mm[i]), int(q[i]), int(g[i]), int(h[i]), float(gc[i])) for i in range(k)]

    def shape_for(self, stage: int, m: int, form: int = FORM_BATCHED):
        q, g, h = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        self._chk(self._L.sx_shape_for(self._h, stage, form, m, ctypes.byref(q), ctypes.byref(g), ctypes.byref(h)))
        return q.value, g.value, h.value
