"""CPU-side checks of the C ABI library: it loads, and exports every entry point that
include/sx_hmmer.h declares (no device calls)."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "sx_hmmer.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(sx_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_the_binding_list(sxa):
    assert declared_symbols() == sorted(sxa.EXPORTS)


def test_library_loads_and_exports_every_declared_symbol(sxa):
    if not os.path.exists(sxa.LIB_PATH):
        pytest.skip("libsxhmmer.so not built")
    lib = ctypes.CDLL(sxa.LIB_PATH)
    for name in declared_symbols():
        assert hasattr(lib, name), name


def test_create_without_device_fails_loudly(sxa):
    """On a host with no HIP device the context refuses to start (no CPU fallback)."""
    if not os.path.exists(sxa.LIB_PATH):
        pytest.skip("libsxhmmer.so not built")
    import torch
    if torch.cuda.is_available():
        pytest.skip("a device is present")
    with pytest.raises(sxa.SxError):
        sxa.Context(0)


class _KernelEntry(ctypes.Structure):
    _fields_ = [("stage", ctypes.c_int), ("q", ctypes.c_int), ("g", ctypes.c_int), ("spec", ctypes.c_int),
                ("hoist", ctypes.c_int), ("wid", ctypes.c_int), ("block", ctypes.c_int), ("lds", ctypes.c_int),
                ("fn", ctypes.c_void_p)]


def _check_aligned_layout(e):
    """Aligned lane blocks (MSV and Viterbi, G = 8 or G >= 16; the MSV bound also G = 4): a
    128-bit LDS read serves 16 lanes a cycle from 64 four-byte banks, so the 16 lanes of
    every quarter-wave must start on 16 distinct four-bank slots whatever residue row each
    group reads.  Viterbi's transition rows (one copy, the same addresses for every group)
    follow."""
    assert e.stage in (0, 1, 4, 5) and e.spec == 0 and (e.g == 8 or e.g >= 16 or (e.g, e.stage) == (4, 4))
    assert e.q % 4 == 0
    # a lane block holds Q floats, or for the int16 bounds (stages 4, 5) Q/2 dwords of two
    # int16 positions each
    per_lane = e.q // 2 if e.stage in (4, 5) else e.q
    assert per_lane % 4 == 0
    if e.g == 4:
        stride = next(v for v in range(per_lane, per_lane + 16) if v % 16 == 4)
    else:
        stride = per_lane + 4 if per_lane % 8 == 0 else per_lane + 8
        assert (stride // 4) % 2 == 1
    copies = {4: 4, 8: 2}.get(e.g, 1)
    skew = 0 if copies == 1 else 64 // copies
    assert e.wid >= e.g * stride
    if copies > 1:
        assert e.wid % 64 == 0  # the residue row moves no bank
    tab = copies * 25 * e.wid + (copies - 1) * skew
    assert e.lds == 4 * (tab + (8 * e.g * stride if e.stage in (1, 5) else 0))
    assert tab % 4 == 0 and (e.g * stride) % 4 == 0  # transition rows stay 16-byte aligned
    copy_stride = 25 * e.wid + skew
    for quarter in range(4):
        slots = set()
        for t in range(16):
            lane = quarter * 16 + t
            grp, j = divmod(lane, e.g)
            c = grp % copies
            # any residue row x adds x * wid, a multiple of 64 when groups share a quarter
            start = c * copy_stride + j * stride
            assert start % 4 == 0
            slots.add((start % 64) // 4)
        assert len(slots) == 16, (e.q, e.g, quarter, sorted(slots))


def _round4(n):
    return (n + 3) // 4 * 4


def _fwd_rec_floats(q, g):
    """Forward's LDS records per lane group: 8 Q + 4 floats of transitions, then Q + 4 or
    Q + 8 floats of D-chain products (the stride whose dword count over 4 is odd)."""
    pstride = q + 4 if ((q + 4) // 4) % 2 else q + 8
    return g * (8 * q + 4) + g * pstride


def test_kernel_table_names_each_shape_once(sxa):
    """Each (stage, Q, G, special rows, table mode) is one distinct kernel; the special-row
    writers are separate kernels from the score-only ones (a launch must never pick one for
    the other)."""
    if not os.path.exists(sxa.LIB_PATH):
        pytest.skip("libsxhmmer.so not built")
    lib = ctypes.CDLL(sxa.LIB_PATH)
    lib.sx_kernel_table.restype = ctypes.POINTER(_KernelEntry)
    n = ctypes.c_int()
    tab = lib.sx_kernel_table(ctypes.byref(n))
    rows = [(tab[i].stage, tab[i].q, tab[i].g, tab[i].spec, tab[i].hoist) for i in range(n.value)]
    fns = [tab[i].fn for i in range(n.value)]
    assert len(set(rows)) == len(rows)
    assert len(set(fns)) == len(fns)
    assert all(r[4] in (0, 1, 2, 3, 4) and r[3] in (0, 1) for r in rows)
    for i in range(n.value):
        e = tab[i]
        # 512 threads also for the long-model Forward scorers of 40 and 48 positions a lane
        # and the Backward of 40 (eight pairs share the staged table: two waves a SIMD)
        wide = e.hoist == 2 and e.spec == 0 and ((e.stage == 2 and e.q in (40, 48)) or (e.stage == 3 and e.q == 40))
        assert e.block == (512 if wide else {3: 512, 4: 1024}.get(e.hoist, 256))
        if e.hoist in (3, 4):
            _check_aligned_layout(e)
        elif e.stage == 2:
            # Forward: emission rows (none for long models), then per lane group the
            # 32-byte transition records of Q positions and the D-chain products
            assert e.lds == 4 * (_round4(0 if e.hoist == 2 else 25 * e.wid) + _fwd_rec_floats(e.q, e.g))
        elif e.hoist == 2:
            assert e.lds == (0 if e.stage == 0 else 4 * 8 * e.wid)
        elif e.stage == 5:
            # one-lane Viterbi bound: odd emission rows, then 16-byte-aligned transition rows
            tw = (e.q // 2 + 3) // 4 * 4
            assert e.lds == 4 * ((25 * e.wid + 3) // 4 * 4 + 8 * tw)
        else:
            assert e.lds == 4 * (25 if e.stage in (0, 4) else 33) * e.wid
        if e.stage in (4, 5):
            # the int16 bound's one-lane rows are odd (distinct banks for distinct rows)
            assert e.hoist in (0, 3, 4) and e.spec == 0
            if e.hoist == 0:
                assert e.g == 1 and e.wid % 2 == 1 and e.wid >= e.q // 2
    long_rows = {r for r in rows if r[4] == 2}
    for st in (0, 1, 2, 3):
        for q in (24, 32, 40, 48, 64):
            assert (st, q, 64, 0, 2) in long_rows
            if st >= 2:
                assert (st, q, 64, 1, 2) in long_rows


def test_encode_matches_the_reference_vectors_and_the_oracle(sxa, orc):
    """sx_encode restates alphabet::encode (alphabet.rs:58-93); the expected codes are the
    reference's own test vectors (alphabet.rs:104-116) and every byte value against the
    oracle's restatement."""
    if not os.path.exists(sxa.LIB_PATH):
        pytest.skip("libsxhmmer.so not built")
    assert list(sxa.encode(b"ACDEFGHIKLMNPQRSTVWY")) == list(range(20))
    assert list(sxa.encode(b"uo")) == [1, 8]
    assert list(sxa.encode(b"BJZX*")) == [20, 21, 22, 23, 24]
    assert list(sxa.encode(b"m k\n")) == [10, 8]
    assert list(sxa.encode(b"-._~!1")) == [24] * 6
    assert list(sxa.encode(b"zq")) == [22, 13]
    assert len(sxa.encode(b"")) == 0 and len(sxa.encode(b" \t\r\n\f")) == 0
    every = bytes(range(256)) * 3
    assert np.array_equal(sxa.encode(every), orc.encode(every))
    # \v is not ASCII whitespace in Rust's is_ascii_whitespace: it is encoded, as '*'
    assert list(sxa.encode(b"\x0bA")) == [24, 0]
